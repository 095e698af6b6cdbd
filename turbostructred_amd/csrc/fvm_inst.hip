// fvm_inst.hip -- instantiates the stage kernels of one (dimension, flux) pair.  Compiled once
// per pair with -DTS_DIM=<1|2|3> -DTS_FLUX=<0 Godunov|1 Roe|3 HLLC> so that the kernel families
// build in parallel.
#include "fvm_dispatch.hpp"
#include "fvm_viscous.hpp"
#if TS_DIM == 3 && __has_include("fvm_kernels3d.hpp")
#include "fvm_kernels3d.hpp"
#define TS_HAVE_3D 1
#endif

#ifndef TS_DIM
#error "TS_DIM must be defined"
#endif
#ifndef TS_FLUX
#error "TS_FLUX must be defined"
#endif

#define TS_CAT2(a, b, c, d) a##b##c##d
#define TS_CAT(a, b, c, d) TS_CAT2(a, b, c, d)
#define TS_GETTER TS_CAT(get_kernels_d, TS_DIM, _f, TS_FLUX)

namespace tsd {

#if !(TS_DIM == 3 && !defined(TS_HAVE_3D))
template <int RC, int LM, int MODE, typename R, typename TIn, typename TOut>
static void launch(const KArgs& a, dim3 g, dim3 b, hipStream_t s) {
#if TS_DIM == 3
    hipLaunchKernelGGL((fvm_stage3d<RC, LM, TS_FLUX, MODE, R, TIn, TOut>), g, b, 0, s, a);
#else
    hipLaunchKernelGGL((fvm_stage<TS_DIM, RC, LM, TS_FLUX, MODE, R, TIn, TOut>), g, b, 0, s, a);
#endif
    if constexpr (ModeTraits<MODE>::sr) hipLaunchKernelGGL(dt_finalize_kernel<0>, dim3(1), dim3(1024), 0, s, a.ctl, a.g.cfl, a.dt_part, a.ntiles);
}

// resident workgroups per CU of the single-solve stage kernels (the minimum over the kernels of
// one precision): sets how many tiles run in one round
template <typename K> static int blocks_per_cu(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}
template <int RC, int LM, typename R, typename S, typename E>
static int single_occupancy() {
#if TS_DIM == 3
    const int T = 64 * ZW;
#define TS_K(MODE, A, B, C) fvm_stage3d<RC, LM, TS_FLUX, MODE, A, B, C>
#else
    const int T = 256;
#define TS_K(MODE, A, B, C) fvm_stage<TS_DIM, RC, LM, TS_FLUX, MODE, A, B, C>
#endif
    int m = blocks_per_cu(TS_K(MODE_F1, R, double, S), T);
    int v[] = {blocks_per_cu(TS_K(MODE_F2, R, double, S), T), blocks_per_cu(TS_K(MODE_S3, R, S, double), T),
               blocks_per_cu(TS_K(MODE_DT, R, double, double), T), blocks_per_cu(TS_K(MODE_S1, R, double, E), T)};
#undef TS_K
    for (int x : v) m = x < m ? x : m;
    return m;
}

#if TS_DIM != 3
template <int RC, int LM, int MODE, typename R, typename TIn, typename TOut>
static void launch_batched(const KArgs* d, const int2* t, const int* ntiles, double* part, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((fvm_stage_batched<TS_DIM, RC, LM, TS_FLUX, MODE, R, TIn, TOut>), dim3(nblk), dim3(256), 0, s,
                       d, t, ntiles, part);
}
template <int RC, int LM, typename R, typename S>
static int batched_occupancy() {
    int a = 0, b = 0, c = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, fvm_stage_batched<TS_DIM, RC, LM, TS_FLUX, MODE_F1, R, double, S>, 256, 0) != hipSuccess) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fvm_stage_batched<TS_DIM, RC, LM, TS_FLUX, MODE_F2, R, double, S>, 256, 0) != hipSuccess) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, fvm_stage_batched<TS_DIM, RC, LM, TS_FLUX, MODE_S3, R, S, double>, 256, 0) != hipSuccess) return 1;
    int m = a < b ? a : b;
    m = m < c ? m : c;
    return m > 0 ? m : 1;
}
#endif

// viscous pre-pass: gradients of the interior cells, then the faces of every axis
static void launch_vis_prepass(const KArgs& a, dim3, dim3, hipStream_t s) {
    const unsigned nb = unsigned((a.plane + 255) / 256);
    hipLaunchKernelGGL(vis_gradients<TS_DIM>, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL((vis_face_flux<TS_DIM, 0>), dim3((a.v.nface[0] + 255) / 256), dim3(256), 0, s, a);
#if TS_DIM > 1
    hipLaunchKernelGGL((vis_face_flux<TS_DIM, 1>), dim3((a.v.nface[1] + 255) / 256), dim3(256), 0, s, a);
#endif
#if TS_DIM > 2
    hipLaunchKernelGGL((vis_face_flux<TS_DIM, 2>), dim3((a.v.nface[2] + 255) / 256), dim3(256), 0, s, a);
#endif
}
template <int RC, int LM, int MODE>
static void launch_vis(const KArgs& a, dim3 g, dim3 b, hipStream_t s) {
#if TS_DIM == 3
    hipLaunchKernelGGL((fvm_stage3d<RC, LM, TS_FLUX, MODE, double, double, double, 1>), g, b, 0, s, a);
#else
    hipLaunchKernelGGL((fvm_stage<TS_DIM, RC, LM, TS_FLUX, MODE, double, double, double, 1>), g, b, 0, s, a);
#endif
    if constexpr (ModeTraits<MODE>::sr) hipLaunchKernelGGL(dt_finalize_kernel<0>, dim3(1), dim3(1024), 0, s, a.ctl, a.g.cfl, a.dt_part, a.ntiles);
}

#if TS_DIM == 1
template <int RC, int LM>
static void launch_euler1d_chunk(const KArgs& a, int steps, hipStream_t s) {
    KArgs b = a;  // the other direction of the ping-pong
    b.uin = a.uout;
    b.uout = const_cast<void*>(a.uin);
    if (a.g.n[0] <= WG1D_LDS_CELLS)
        hipLaunchKernelGGL((fvm_euler1d_workgroup<RC, LM, TS_FLUX, true>), dim3(1), dim3(64 * wg1d_waves<TS_FLUX>), 0, s,
                           a, b, steps);
    else
        hipLaunchKernelGGL((fvm_euler1d_workgroup<RC, LM, TS_FLUX, false>), dim3(1), dim3(64 * wg1d_waves<TS_FLUX>), 0, s,
                           a, b, steps);
}
#endif

template <int RC, int LM>
static void fill(int precision, KernelSet* ks) {
    if (precision == 0) {
        ks->dt = launch<RC, LM, MODE_DT, double, double, double>;
        ks->s1e = launch<RC, LM, MODE_S1, double, double, double>;
        ks->s1r = ks->s1e;
        ks->s2 = launch<RC, LM, MODE_S2, double, double, double>;
        ks->s3 = launch<RC, LM, MODE_S3, double, double, double>;
        ks->f1 = launch<RC, LM, MODE_F1, double, double, double>;
        ks->f2 = launch<RC, LM, MODE_F2, double, double, double>;
        ks->rn = launch<RC, LM, MODE_RN, double, double, double>;
        ks->socc = single_occupancy<RC, LM, double, double, double>;
#if TS_DIM != 3
        ks->bf1 = launch_batched<RC, LM, MODE_F1, double, double, double>;
        ks->bf2 = launch_batched<RC, LM, MODE_F2, double, double, double>;
        ks->bs3 = launch_batched<RC, LM, MODE_S3, double, double, double>;
        ks->bocc = batched_occupancy<RC, LM, double, double>;
#endif
        ks->vpre = launch_vis_prepass;
        ks->vdt = launch_vis<RC, LM, MODE_DT>;
        ks->vs1 = launch_vis<RC, LM, MODE_S1>;
        ks->vs2 = launch_vis<RC, LM, MODE_S2>;
        ks->vs3 = launch_vis<RC, LM, MODE_S3>;
        ks->vrn = launch_vis<RC, LM, MODE_RN>;
#if TS_DIM == 1
        ks->euler1d_chunk = launch_euler1d_chunk<RC, LM>;
#endif
    } else {
        ks->dt = launch<RC, LM, MODE_DT, float, double, double>;
        ks->s1e = launch<RC, LM, MODE_S1, float, double, double>;
        ks->s1r = launch<RC, LM, MODE_S1, float, double, float>;
        ks->s2 = launch<RC, LM, MODE_S2, float, float, float>;
        ks->s3 = launch<RC, LM, MODE_S3, float, float, double>;
        ks->f1 = launch<RC, LM, MODE_F1, float, double, float>;
        ks->f2 = launch<RC, LM, MODE_F2, float, double, float>;
        ks->rn = launch<RC, LM, MODE_RN, float, double, float>;
        ks->socc = single_occupancy<RC, LM, float, float, double>;
#if TS_DIM != 3
        ks->bf1 = launch_batched<RC, LM, MODE_F1, float, double, float>;
        ks->bf2 = launch_batched<RC, LM, MODE_F2, float, double, float>;
        ks->bs3 = launch_batched<RC, LM, MODE_S3, float, float, double>;
        ks->bocc = batched_occupancy<RC, LM, float, float>;
#endif
    }
}
#endif

#if TS_DIM == 3 && !defined(TS_HAVE_3D)
bool TS_GETTER(int, int, KernelSet*) { return false; }
#else
bool TS_GETTER(int rv, int precision, KernelSet* ks) {
    switch (rv) {
    case 0: fill<RC_PWC, LIM_BJ>(precision, ks); return true;
    case 1: fill<RC_LINEAR, LIM_BJ>(precision, ks); return true;
    case 2: fill<RC_ENO2, LIM_BJ>(precision, ks); return true;
    case 3: fill<RC_LINLIM, LIM_BJ>(precision, ks); return true;
    case 4: fill<RC_LINLIM, LIM_VENKAT>(precision, ks); return true;
    }
    return false;
}
#endif

}  // namespace tsd
