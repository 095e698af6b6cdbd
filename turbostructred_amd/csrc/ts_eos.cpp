// ts_eos.cpp -- ts_eos_eval: the device equation-of-state functor evaluated over host arrays.
#include <hip/hip_runtime.h>

#include "ts_internal.hpp"

namespace {

template <int M>
__global__ void __launch_bounds__(256) eos_kernel(double gamma, const double* __restrict__ ro,
                                                  const double* __restrict__ x, size_t n, double* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const tsd::IdealGasEOS<double> eos(gamma);
    double r;
    if constexpr (M == 0) r = eos.GetPressure(ro[i], x[i]);
    else if constexpr (M == 1) r = eos.GetSoundSpeed(ro[i], x[i]);
    else if constexpr (M == 2) r = eos.GetInternalEnergy(ro[i], x[i]);
    else if constexpr (M == 3) r = eos.GetPressureEnergyDerivative(ro[i], x[i]);
    else r = eos.GetPressureDensityDerivative(ro[i], x[i]);
    out[i] = r;
}

}  // namespace

extern "C" int ts_eos_eval(double gamma, int method, const double* ro, const double* x, size_t n, double* out,
                           int device) {
    if (method < 0 || method > 4 || (n && (!ro || !x || !out))) {
        tsr::set_thread_error("ts_eos_eval: bad method or null array");
        return TS_ERR_INVALID;
    }
    if (n == 0) return TS_OK;
    double* d = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc((void**)&d, 3 * n * sizeof(double));
    if (e != hipSuccess) {
        tsr::set_thread_error("ts_eos_eval: " + tsr::hip_msg(e, "hipMalloc"));
        return TS_ERR_HIP;
    }
    double *dr = d, *dx = d + n, *dout = d + 2 * n;
    e = hipMemcpy(dr, ro, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const dim3 grid(unsigned((n + 255) / 256)), block(256);
        switch (method) {
        case 0: hipLaunchKernelGGL(eos_kernel<0>, grid, block, 0, nullptr, gamma, dr, dx, n, dout); break;
        case 1: hipLaunchKernelGGL(eos_kernel<1>, grid, block, 0, nullptr, gamma, dr, dx, n, dout); break;
        case 2: hipLaunchKernelGGL(eos_kernel<2>, grid, block, 0, nullptr, gamma, dr, dx, n, dout); break;
        case 3: hipLaunchKernelGGL(eos_kernel<3>, grid, block, 0, nullptr, gamma, dr, dx, n, dout); break;
        default: hipLaunchKernelGGL(eos_kernel<4>, grid, block, 0, nullptr, gamma, dr, dx, n, dout); break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        tsr::set_thread_error("ts_eos_eval: " + tsr::hip_msg(e, "device evaluation"));
        return TS_ERR_HIP;
    }
    return TS_OK;
}
