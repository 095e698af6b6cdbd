// ts_layout.cpp -- conversion between the reference's AoS state layout (5 values per cell, ghost
// layers included, kernel.h:93-96) and the library's SoA interior planes, on the device: the host
// only copies the caller's array as it is.
#include <hip/hip_runtime.h>

#include "ts_internal.hpp"

namespace {

struct Shape {
    int n[3];           // interior cells per axis
    int g[3];           // ghost layers per axis (1 on active axes)
    size_t plane;       // cells per SoA plane
};

// one thread per interior cell
__device__ __forceinline__ size_t aos_cell(const Shape& s, unsigned c, int& i, int& j, int& k) {
    i = int(c % unsigned(s.n[0]));
    const unsigned r = c / unsigned(s.n[0]);
    j = int(r % unsigned(s.n[1]));
    k = int(r / unsigned(s.n[1]));
    const size_t nxa = size_t(s.n[0]) + 2 * s.g[0], nya = size_t(s.n[1]) + 2 * s.g[1];
    return ((size_t(k + s.g[2]) * nya + size_t(j + s.g[1])) * nxa + size_t(i + s.g[0])) * 5;
}

__global__ void __launch_bounds__(256) aos_to_soa_kernel(const double* __restrict__ aos, double* __restrict__ soa,
                                                         Shape s) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= s.plane) return;
    int i, j, k;
    const size_t a = aos_cell(s, unsigned(c), i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) soa[size_t(v) * s.plane + c] = aos[a + v];
}

__global__ void __launch_bounds__(256) soa_to_aos_kernel(const double* __restrict__ soa, double* __restrict__ aos,
                                                         Shape s) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= s.plane) return;
    int i, j, k;
    const size_t a = aos_cell(s, unsigned(c), i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) aos[a + v] = soa[size_t(v) * s.plane + c];
}

Shape shape_of(int nd, const int n[3]) {
    Shape s;
    for (int a = 0; a < 3; a++) {
        s.n[a] = n[a];
        s.g[a] = a < nd ? 1 : 0;
    }
    s.plane = size_t(n[0]) * n[1] * n[2];
    return s;
}

}  // namespace

namespace tsr {

size_t aos_doubles(int nd, const int n[3]) {
    size_t t = 5;
    for (int a = 0; a < 3; a++) t *= size_t(n[a]) + (a < nd ? 2 : 0);
    return t;
}

hipError_t device_aos_to_soa(const double* d_aos, double* d_soa, int nd, const int n[3], hipStream_t stream) {
    const Shape s = shape_of(nd, n);
    if (s.plane == 0) return hipSuccess;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(unsigned((s.plane + 255) / 256)), dim3(256), 0, stream, d_aos, d_soa, s);
    return hipGetLastError();
}

hipError_t device_soa_to_aos(const double* d_soa, double* d_aos, int nd, const int n[3], hipStream_t stream) {
    const Shape s = shape_of(nd, n);
    if (s.plane == 0) return hipSuccess;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(unsigned((s.plane + 255) / 256)), dim3(256), 0, stream, d_soa, d_aos, s);
    return hipGetLastError();
}

}  // namespace tsr
