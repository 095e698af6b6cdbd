// fvm_dispatch.hpp -- table of the compiled kernel instances, filled per (dimension, flux)
// translation unit and looked up by the host runtime at ts_create.
#pragma once

#include "fvm_kernels.hpp"

namespace tsd {

using LaunchFn = void (*)(const KArgs&, dim3, dim3, hipStream_t);
// batched (persistent) launch: descriptors, (problem, tile) table, device tile count, per-entry
// minima of V/SR (written by stage 1), number of 256-thread blocks
using BatchLaunchFn = void (*)(const KArgs*, const int2*, const int*, double*, int, hipStream_t);
using OccupancyFn = int (*)();  // resident workgroups per CU

// dt pass, Euler stage (double out), SSP stage 1 (stage-array out), SSP stages 2 and 3, and the
// fused-dt SSP stages 1 and 2
struct KernelSet {
    LaunchFn dt = nullptr, s1e = nullptr, s1r = nullptr, s2 = nullptr, s3 = nullptr;
    LaunchFn f1 = nullptr, f2 = nullptr;
    LaunchFn rn = nullptr;  // residual norms of the stencil input (no state written)
    BatchLaunchFn bf1 = nullptr, bf2 = nullptr, bs3 = nullptr;  // fused-dt SSP-RK3, batched
    OccupancyFn bocc = nullptr;
    OccupancyFn socc = nullptr;  // resident workgroups per CU of the single-solve kernels
    // FP64 instances with viscous terms and body forces: the pre-pass (velocity
    // gradients and viscous face fluxes of the stage input), the dt pass, the three stages and
    // the residual norms
    LaunchFn vpre = nullptr, vdt = nullptr, vs1 = nullptr, vs2 = nullptr, vs3 = nullptr, vrn = nullptr;
    // small 1-D forward-Euler solves: a chunk of steps in one workgroup (KArgs.uin = U^n,
    // uout = the other state buffer); FP64, 1-D units only
    using ChunkFn = void (*)(const KArgs&, int steps, hipStream_t);
    ChunkFn euler1d_chunk = nullptr;
};

// reconstruction variants: 0 PWC, 1 linear, 2 ENO2, 3 limited linear BJ, 4 limited linear Venkat
bool get_kernels_d1_f0(int rv, int precision, KernelSet* ks);
bool get_kernels_d1_f1(int rv, int precision, KernelSet* ks);
bool get_kernels_d1_f3(int rv, int precision, KernelSet* ks);
bool get_kernels_d2_f0(int rv, int precision, KernelSet* ks);
bool get_kernels_d2_f1(int rv, int precision, KernelSet* ks);
bool get_kernels_d2_f3(int rv, int precision, KernelSet* ks);
bool get_kernels_d3_f0(int rv, int precision, KernelSet* ks);
bool get_kernels_d3_f1(int rv, int precision, KernelSet* ks);
bool get_kernels_d3_f3(int rv, int precision, KernelSet* ks);

}  // namespace tsd
