// fvm_inst.hip -- instantiates the stage kernels of one (dimension, flux) pair.  Compiled once
// per pair with -DTS_DIM=<1|2|3> -DTS_FLUX=<0 Godunov|1 Roe|3 HLLC> so that the kernel families
// build in parallel.
#include "fvm_dispatch.hpp"
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
}

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
    } else {
        ks->dt = launch<RC, LM, MODE_DT, float, double, double>;
        ks->s1e = launch<RC, LM, MODE_S1, float, double, double>;
        ks->s1r = launch<RC, LM, MODE_S1, float, double, float>;
        ks->s2 = launch<RC, LM, MODE_S2, float, float, float>;
        ks->s3 = launch<RC, LM, MODE_S3, float, float, double>;
        ks->f1 = launch<RC, LM, MODE_F1, float, double, float>;
        ks->f2 = launch<RC, LM, MODE_F2, float, double, float>;
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
