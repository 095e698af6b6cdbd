// fvm_kernels.hpp -- fused finite-volume stage kernels for gfx950 (1-D and 2-D marching form).
//
// One wavefront owns a strip of 64 columns along x (60 output columns plus a halo of two
// lanes on each side) and marches along y over `ty` output rows.  In 2-D it keeps a three-row
// window of the state, and the values carried from one row to the next, in its own LDS slots
// (no other lane reads them, so there are no barriers), with the next row in flight in
// registers (prefetched one iteration ahead).  In 2-D, x-neighbours are read from the same LDS
// rows at the adjacent lanes (1-D and 3-D use DPP wave shifts); every face flux is evaluated
// once, by the lane on its plus side, and the residual of a cell is accumulated from its x
// faces, then its y faces (ERK:1232,1312) before the stage update is applied and the row is
// stored with coalesced SoA writes.
//
// Modes
//   MODE_DT  spectral radius of every cell from the face eigenvalues, min(V/SR) over the grid
//            (wave shuffle, one stored minimum per tile); a one-workgroup kernel launched
//            after it turns the minima into dt (ERK:1427-1437).  Used before forward Euler.
//   MODE_S1  U1 = U + dt L(U)          forward Euler (kernel.h:319-331) / SSP stage 1
//   MODE_S2  U2 = 3/4 Un + 1/4 (U1 + dt L(U1))
//   MODE_S3  U  = 1/3 Un + 2/3 (U2 + dt L(U2))   (written in place over Un: Un is read
//            point-wise by the thread that overwrites it)
//   MODE_F1  SSP stage 1 with the dt pass fused in: writes L(Un) = R/V and does the dt
//            reduction of MODE_DT from the same face eigenvalues.
//   MODE_F2  SSP stage 2 reading Un and L(Un): its stencil values are U1 = Un + dt L(Un)
//            formed on the fly, so U1 never goes to HBM.
//   MODE_RN  residual norms sum(V |R|) per variable of UpdateSolution (kernel.h:326) over the
//            interior, writing nothing else: per-block partial sums, and the last block adds
//            them in block order (a deterministic sum).  Launched after the dt pass of every
//            step when the norms are asked for, it runs only in the step that ends the
//            ts_step call (the norms reported), so a call costs one extra residual pass.
#pragma once

#include <type_traits>

#include "fvm_device.hpp"

namespace tsd {

enum : int { MODE_DT = 0, MODE_S1 = 1, MODE_S2 = 2, MODE_S3 = 3, MODE_F1 = 4, MODE_F2 = 5, MODE_RN = 6 };
constexpr int STRIP = 60;   // output columns per wavefront
constexpr int ZW = 8;         // wavefronts (z-planes) per 3-D workgroup
constexpr int ZOUT = ZW - 2;  // output z-planes per 3-D workgroup
constexpr int MAX_BC = 8;

// Device-resident control block of one solve: the step bookkeeping of StepInfo (kernel.h:56) and
// the termination rule of Kernel::Run (kernel.h:1834-1846), evaluated on the device so that a
// chunk of steps can be issued without host round trips.
struct StepCtl {
    double time, dt, t_max, next_snap;
    long long iter;
    int max_steps, steps_done;
    int active;  // 1 while steps remain to be taken
    int go;      // 1 if the current step runs (set by the dt pass)
    int error;   // 1 = dt not finite and positive
    int pad0;
    double* dt_log;
    long long dt_log_base, dt_log_cap;  // dt of iteration `it` goes to dt_log[it - base]
    double rnorm[5];                    // residual norms of the last step (MODE_RN)
    unsigned int rn_done, pad2;
};

struct GridArgs {
    int nd;
    int n[3];
    int periodic[3];
    const double* h[3];   // cell sizes, ghost layer included (n + 2 entries)
    const double* rc[3];  // 1/(x_{i+1} - x_{i-1}), interior cells
    const double* rl[3];  // 1/(x_i - x_{i-1})
    const double* rr[3];  // 1/(x_{i+1} - x_i)
    const double* ih[3];  // 1/h, ghost layer included
    double eps2;          // Venkatakrishnan eps^2 of every cell (uniform grids)
    // non-uniform grids: eps^2 of cell (i, j, k) = ev[0][i] * ev[1][j] * ev[2][k], the per-axis
    // factors (K h_a)^(3/nd) of (K (hx hy hz)^(1/nd))^3 (Limiters.h:93-96); null when uniform
    const double* ev[3];
    double gamma, eps, cfl;
    BCDev bc[MAX_BC];
    int side_marker[6];
    const uint8_t* face_marker[6];
};

// Viscous terms and body forces (ERK:79-874, kernel.h:941-989), read only by the VIS instances
// of the stage kernels and by the viscous pre-pass.
struct VisArgs {
    const double* x[3];  // cell-centre coordinates, ghost layer included (n + 2 entries)
    double* grad;        // velocity gradients of the interior cells: plane d*nd + e = d vel_d / d x_e
    double* fv[3];       // viscous fluxes of the faces normal to axis a: 4 planes (momentum x, y, z,
                         // energy), faces numbered like cells with n_a + 1 along a
    unsigned nface[3];   // faces normal to axis a (the plane size of fv[a])
    double mu;           // gas viscosity (viscous fluxes and the spectral-radius term)
    double sigma[3];     // potential force (IsExternalForceRequared)
    double accel[3];     // uniform acceleration (IsUnifromAccelerationRequared)
    int viscous, force, accel_on;
    int gate_go;         // pre-pass: runs when the step runs (1) or while steps remain (0)
};

struct KArgs {
    GridArgs g;
    VisArgs v;
    const void* uin;   // stencil input planes (TIn)
    const void* lin;   // L(Un) planes (TOut type), MODE_F2
    const double* un;  // U^n planes (SSP stages 2, 3)
    void* uout;        // output planes (TOut)
    void* dout;        // FP32-residual stage 2: the increment U2 - U^n (float planes), or null
    const void* din;   // FP32-residual stage 3: that increment (float planes), or null
    size_t plane;      // cells per plane
    StepCtl* ctl;
    int ty;            // output rows per tile
    int nstrips;       // strips along x
    int ntiles;        // strips * row blocks
    double* rn_part;   // per-block partial residual norms, 5 per block (MODE_RN)
    double* dt_part;   // single solves: each tile's minimum of V/SR (dt passes, fused stage 1)
};

// ---- wave-level helpers -------------------------------------------------------------------------
// lane i receives lane i-1 (wave_shr:1) / lane i+1 (wave_shl:1); the end lanes are halo lanes
// and receive 0 (bound_ctrl), which needs no copy of an "old" value before each move.
__device__ inline double dpp_from_left(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline double dpp_from_right(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x130, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline float dpp_from_left(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ inline float dpp_from_right(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// tiles so that neighbouring tiles (which share halo rows/columns) share an L2.
__device__ inline int xcd_remap(int b, int nb) {
    int q = nb >> 3, r = nb & 7;
    int x = b & 7, l = b >> 3;
    return x * q + (x < r ? x : r) + l;
}

// Map a possibly out-of-range index to its source cell: periodic wrap, or the mirror image
// of the ghost fill (ghost i = -l takes cell l-1, kernel.h:734-735, 751-752).
__device__ inline void map_index(int x, int n, int periodic, int& xs, int& side) {
    side = 0;
    if (periodic) {
        xs = x % n;
        if (xs < 0) xs += n;
        return;
    }
    if (x < 0) {
        xs = -x - 1;
        side = -1;
    } else if (x >= n) {
        xs = 2 * n - 1 - x;
        side = 1;
    } else {
        xs = x;
    }
    xs = xs < 0 ? 0 : (xs >= n ? n - 1 : xs);
}

__device__ inline int bc_marker(const GridArgs& g, int side, int face_index) {
    const uint8_t* fm = g.face_marker[side];
    return fm ? int(fm[face_index]) : g.side_marker[side];
}

// Variable v of an SoA state array at a cell: a wave-uniform 64-bit plane base plus a 32-bit
// per-lane byte offset -- the global load/store form with a scalar base address, so that each
// access needs one VGPR of address and no 64-bit address arithmetic.  A plane stays below
// 4 GiB (checked when a context is created).  (Casting the base to the global address space
// explicitly changed nothing for the 1-D / 2-D stages and made the viscous stages 10-25% slower.)
template <typename T>
__device__ __forceinline__ T ld_plane(const T* base, size_t P, int v, unsigned cell) {
    const char* b = reinterpret_cast<const char*>(base + (size_t)v * P);
    return *reinterpret_cast<const T*>(b + cell * unsigned(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st_plane(T* base, size_t P, int v, unsigned cell, T x) {
    char* b = reinterpret_cast<char*>(base + (size_t)v * P);
    *reinterpret_cast<T*>(b + cell * unsigned(sizeof(T))) = x;
}

// The same store with the non-temporal hint, for the 1-D / 2-D stage outputs: each value is
// written once and is next read by the following launch, so keeping it out of the way of the
// stencil rows still being re-read from L2 pays (about 3 % per 2-D step in FP64); the 3-D kernel
// keeps the plain store, which measured faster there.
template <typename T>
__device__ __forceinline__ void st_plane_stream(T* base, size_t P, int v, unsigned cell, T x) {
    char* b = reinterpret_cast<char*>(base + (size_t)v * P);
    __builtin_nontemporal_store(x, reinterpret_cast<T*>(b + cell * unsigned(sizeof(T))));
}

// The batched kernel reads its descriptors, and so every state and geometry pointer, from memory:
// as generic pointers their accesses became flat instructions (which also wait on the scalar
// loads' counter) and the wave-uniform row geometry came from vector loads, whose waits also
// waited for the stencil prefetch.  Its instances (GP) address the state planes and per-lane
// geometry as global memory and the row geometry, read-only for the whole run, as constant
// memory (scalar loads).  The single-solve kernels take their pointers from kernel arguments,
// where the compiler infers as much, and keep the generic forms (the explicit casts slowed the
// viscous stages).
#define TS_AS_GLOBAL __attribute__((address_space(1)))
#define TS_AS_CONST __attribute__((address_space(4)))
template <bool GP, typename T>
__device__ __forceinline__ T ld_state(const T* base, size_t P, int v, unsigned cell) {
    if constexpr (GP) {
        const TS_AS_GLOBAL char* b = (const TS_AS_GLOBAL char*)(base + (size_t)v * P);
        return *(const TS_AS_GLOBAL T*)(b + cell * unsigned(sizeof(T)));
    } else {
        return ld_plane(base, P, v, cell);
    }
}
template <bool GP, typename T>
__device__ __forceinline__ void st_state_stream(T* base, size_t P, int v, unsigned cell, T x) {
    if constexpr (GP) {
        TS_AS_GLOBAL char* b = (TS_AS_GLOBAL char*)(base + (size_t)v * P);
        __builtin_nontemporal_store(x, (TS_AS_GLOBAL T*)(b + cell * unsigned(sizeof(T))));
    } else {
        st_plane_stream(base, P, v, cell, x);
    }
}
// per-lane geometry
template <bool GP, typename T>
__device__ __forceinline__ T ld_geo(const T* p, int i) {
    if constexpr (GP) return ((const TS_AS_GLOBAL T*)p)[i];
    else return p[i];
}
// wave-uniform geometry (row sizes and slope factors)
template <bool GP, typename T>
__device__ __forceinline__ T ld_geo_uniform(const T* p, int i) {
    if constexpr (GP) return ((const TS_AS_CONST T*)p)[i];
    else return p[i];
}


// V/SR of one cell; a NaN or negative spectral radius (non-physical state) becomes 0 so that
// the step fails with dt = 0 instead of silently skipping the cell in the minimum.
__device__ inline double dt_candidate(double vol, double sr) {
    double q = vol / sr;
    return (q > 0.0) ? q : 0.0;
}

// dt = min(V/SR) * CFL (ERK:1430) from the grid minimum mn, clamped to the next snapshot
// (ERK:1433); advances Time / Iteration, logs dt and evaluates the stop rule of Kernel::Run
// (kernel.h:1834-1846).  One thread.
__device__ inline void dt_apply(StepCtl* ctl, double cfl, double mn) {
    double dtn = mn * cfl;
    double t0 = ctl->time;
    if (ctl->next_snap > 0 && ctl->next_snap < t0 + dtn) dtn = ctl->next_snap - t0;
    if (!(dtn > 0.0) || !(dtn < __builtin_inf())) {
        ctl->error = 1;
        ctl->go = 0;
        ctl->active = 0;
    } else {
        double t1 = t0 + dtn;
        ctl->dt = dtn;
        ctl->time = t1;
        long long it = ctl->iter;
        ctl->iter = it + 1;
        int sd = ctl->steps_done + 1;
        ctl->steps_done = sd;
        long long li = it - ctl->dt_log_base;
        if (ctl->dt_log && li >= 0 && li < ctl->dt_log_cap) ctl->dt_log[li] = dtn;
        ctl->go = 1;
        bool stop = (t1 >= ctl->t_max) || (ctl->next_snap > 0 && t1 == ctl->next_snap) || (sd >= ctl->max_steps);
        ctl->active = stop ? 0 : 1;
    }
}

// Single solves: the step's dt from the `n` per-tile minima of V/SR that a dt pass or fused
// stage 1 stored, launched right after it.  Plain stores and one workgroup's reduction instead of
// one same-address atomic per wavefront: those serialise where the wavefronts finish together,
// at the end of the launch (a 1-D stage 1 with 70k strips spent most of its time in them).  The
// minimum is exact, so its order does not matter.  Skipped when no step is running.  (A template
// only so that the translation units including this header share one definition.)
template <int = 0>
__global__ void __launch_bounds__(1024) dt_finalize_kernel(StepCtl* ctl, double cfl, const double* part, int n) {
    if (!ctl->active) return;  // the same for every thread: before the barrier
    double m = __builtin_inf();
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = __builtin_fmin(m, part[i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = __builtin_fmin(m, __shfl_xor(m, o));
    __shared__ double red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < (blockDim.x >> 6); w++) m = __builtin_fmin(m, red[w]);
        dt_apply(ctl, cfl, m);
    }
}

// Batched solves: the dt of every resident solve from its work items' minima of V/SR, which
// fused stage 1 stored by table entry (a solve's entries are contiguous: `ranges` holds the
// descriptor slot, first entry and count of each member).  One wavefront per member, launched
// right after stage 1; skipped for a solve whose run has ended.  Per-item atomics on the solve's
// control block serialised on a lone large solve (2 x 3000 same-address atomics per stage 1 at
// 2048^2), and their ordering needed either a fence, which writes back the L2, or waiting for a
// returned value.
template <int = 0>
__global__ void __launch_bounds__(256) dt_finalize_batched(const KArgs* __restrict__ descs, const int4* __restrict__ ranges,
                                                           const int* __restrict__ nmem_p, const double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int nm = __builtin_amdgcn_readfirstlane(*nmem_p);
    const int stride = gridDim.x * (blockDim.x >> 6);
    for (int i = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); i < nm; i += stride) {
        const int4 r = ranges[i];  // slot, first entry, entries
        const KArgs& a = descs[r.x];
        StepCtl* ctl = a.ctl;
        if (!ctl->active) continue;
        double m = __builtin_inf();
        for (int j = lane; j < r.z; j += 64) m = __builtin_fmin(m, part[r.y + j]);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = __builtin_fmin(m, __shfl_xor(m, o));
        if (lane == 0) dt_apply(ctl, a.g.cfl, m);
    }
}

// Block sum of the per-lane residual norms (wave shuffle tree, then the waves in order), one
// partial per block, and the last block to finish adds the partials in block order.
__device__ inline void rn_reduce_finalize(StepCtl* ctl, double* part, const double (&rn)[NV], double* red,
                                          int lane, int wave, int nwaves) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double x = rn[v];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
        if (lane == 0) red[wave * NV + v] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 0; v < NV; v++) {
            double x = red[v];
            for (int w = 1; w < nwaves; w++) x += red[w * NV + v];
            part[(size_t)blockIdx.x * NV + v] = x;
        }
        __threadfence();
        unsigned prev = atomicAdd(&ctl->rn_done, 1u);
        if (prev == gridDim.x - 1) {
            __threadfence();
            for (int v = 0; v < NV; v++) {
                double x = 0.0;
                for (unsigned b = 0; b < gridDim.x; b++) x += part[(size_t)b * NV + v];
                ctl->rnorm[v] = x;
            }
            atomicExch(&ctl->rn_done, 0u);
        }
    }
}

template <int MODE> struct ModeTraits {
    static constexpr bool sr = MODE == MODE_DT || MODE == MODE_F1;  // needs V/SR and dt
    static constexpr bool flux = MODE != MODE_DT;                    // needs fluxes
    static constexpr bool use_un = MODE == MODE_S2 || MODE == MODE_S3 || MODE == MODE_F2;
    static constexpr double ca = (MODE == MODE_S3) ? 1.0 / 3.0 : 0.75;
    static constexpr double cb = (MODE == MODE_S3) ? 2.0 / 3.0 : 0.25;
};

// ---- one tile of the 1-D / 2-D stage: a strip of 60 columns marching over ty rows ----------------
// block_min receives the tile's min(V/SR) (sr modes).  `tile` < 0 marks a wavefront without work
// that still has to take part in a block reduction.
// max(4/(3 rho_f), gamma/rho_f) mu fS^2 / V: the viscous part of a cell's spectral radius from
// one face, rho_f = sqrt(rho_L rho_R) of the face states (ERK:1239-1248)
__device__ inline double viscous_sr(double rho_l, double rho_r, double gamma, double mu, double fS, double vol) {
#pragma clang fp contract(off)
    const double ro = __builtin_sqrt(rho_l * rho_r);
    const double a = 4.0 / (3.0 * ro), b = gamma / ro;
    double v = (a < b) ? b : a;
    v *= mu / 1.0;
    v *= fS * fS;
    v /= vol;
    return v;
}

// residual += force and acceleration sources of a cell (kernel.h:941-989)
template <typename R>
__device__ inline void add_sources(const VisArgs& v, const R (&u)[NV], double vol, R (&res)[NV]) {
    const double ro = double(u[0]);
    const double vx = double(u[1]) / ro, vy = double(u[2]) / ro, vz = double(u[3]) / ro;
    if (v.force) {
        const double* s = v.sigma;
        res[1] = R(double(res[1]) + vol * s[0]);
        res[2] = R(double(res[2]) + vol * s[1]);
        res[3] = R(double(res[3]) + vol * s[2]);
        res[4] = R(double(res[4]) + (s[0] * vx + s[1] * vy + s[2] * vz) * vol);
    }
    if (v.accel_on) {
        const double* g = v.accel;
        res[1] = R(double(res[1]) + ro * vol * g[0]);
        res[2] = R(double(res[2]) + ro * vol * g[1]);
        res[3] = R(double(res[3]) + ro * vol * g[2]);
        res[4] = R(double(res[4]) + (ro * g[0] * vx + ro * g[1] * vy + ro * g[2] * vz) * vol);
    }
}

// VIS instances (FP64) add the viscous face fluxes of the pre-pass (subtracted from the convective
// flux before it is differenced, ERK:1221-1222), the viscous spectral-radius term after each
// plus face's eigenvalue, and the body-force sources after the flux sweeps.
// DT_ARG instances take the step's dt from `dt_arg` instead of the control block in memory;
// LDS_STATE instances read and write the state through `uin_arg` / `uout_arg` with plane stride
// `plane_arg` (LDS arrays, addressed generically) with plain stores.
template <int DIM, int RECON, int LIM, int FLUX, int MODE, typename R, typename TIn, typename TOut, int VIS = 0,
          bool DT_ARG = false, bool LDS_STATE = false, bool GP = false>
__device__ __forceinline__ void stage_tile(const KArgs& a, int tile, int lane, double& block_min,
                                           double* rn = nullptr, double dt_arg = 0.0, const void* uin_arg = nullptr,
                                           void* uout_arg = nullptr, size_t plane_arg = 0) {
    using MT = ModeTraits<MODE>;
    const GridArgs& g = a.g;
    const StepCtl* ctl = a.ctl;
    const bool tile_ok = tile >= 0 && tile < a.ntiles;
    const int strip = tile_ok ? tile % a.nstrips : 0;
    const int yb = tile_ok ? tile / a.nstrips : 0;

    const int nx = g.n[0];
    const int ny = (DIM > 1) ? g.n[1] : 1;
    const int xg = strip * STRIP - 2 + lane;
    int xs, xside;
    map_index(xg, nx, g.periodic[0], xs, xside);
    const bool out_lane = tile_ok && lane >= 2 && lane < 2 + STRIP && xg < nx;
    // lane holding the right ghost column: its minus-face flux is the right boundary face
    const bool ghost_x_lane = tile_ok && xg == nx && lane >= 2 && lane <= 62;

    const R gamma = R(g.gamma), eps = R(g.eps);
    const double dt = MT::sr ? 0.0 : (DT_ARG ? dt_arg : ctl->dt);

    CellGeo<DIM, R> geo;
    const double hx = ld_geo<GP>(g.h[0], xs + 1);
    const double ihx = ld_geo<GP>(g.ih[0], xs + 1);
    geo.hh[0] = R(0.5 * hx);
    geo.rc[0] = R(ld_geo<GP>(g.rc[0], xs));
    geo.rl[0] = R(ld_geo<GP>(g.rl[0], xs));
    geo.rr[0] = R(ld_geo<GP>(g.rr[0], xs));
    geo.eps2 = R(g.eps2);
    constexpr bool VENKAT = RECON == RC_LINLIM && LIM == LIM_VENKAT;
    // per-cell eps^2 on non-uniform grids (a wave-uniform branch, Venkatakrishnan instances only)
    const bool ev_cells = VENKAT && g.ev[0] != nullptr;
    const double ev_x = ev_cells ? ld_geo<GP>(g.ev[0], xs) : 0.0;
    if (ev_cells && DIM == 1) geo.eps2 = R(ev_x);

    const TIn* __restrict__ uin = LDS_STATE ? (const TIn*)uin_arg : (const TIn*)a.uin;
    const TOut* __restrict__ lin = (const TOut*)a.lin;
    TOut* uo = LDS_STATE ? (TOut*)uout_arg : (TOut*)a.uout;
    const size_t P = LDS_STATE ? plane_arg : a.plane;
    auto store = [&](int v, unsigned off, TOut x) {
        if constexpr (LDS_STATE) st_plane(uo, P, v, off, x);
        else st_state_stream<GP>(uo, P, v, off, x);
    };

    // One stencil value, the input state at a (possibly ghost) cell of row r, in two steps: the
    // loads (issue_row), and the value formed from them with its ghost image (complete_row).  The
    // prefetched row is completed one iteration after its loads are issued, so that no arithmetic
    // on the loaded values (stage 2's U^n + dt L, a ghost image) makes the wavefront wait for them
    // early.
    struct RawRow {
        TIn u[NV];
        TOut l[MODE == MODE_F2 ? NV : 1];
        int rs, yside;
    };
    auto issue_row = [&](int r, RawRow& q) {
        int rs = 0, yside = 0;
        if constexpr (DIM > 1) map_index(r, ny, g.periodic[1], rs, yside);
        q.rs = rs;
        q.yside = yside;
        const unsigned off = unsigned(rs) * unsigned(nx) + unsigned(xs);
#pragma unroll
        for (int v = 0; v < NV; v++) q.u[v] = ld_state<GP>(uin, P, v, off);
        if constexpr (MODE == MODE_F2) {
#pragma unroll
            for (int v = 0; v < NV; v++) q.l[v] = ld_state<GP>(lin, P, v, off);
        }
    };
    auto complete_row = [&](const RawRow& q, R (&u)[NV]) {
        if constexpr (MODE == MODE_F2) {
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = R(double(q.u[v]) + dt * double(q.l[v]));
        } else {
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = R(q.u[v]);
        }
        if (xside != 0) {
            // one branch per side, so that the side (and, along a row, the condition) is
            // wave-uniform and its data come from scalar loads
            R t[NV];
            if (xside < 0) bc_image(g.bc[bc_marker(g, 0, q.rs)], gamma, 0, u, t);
            else bc_image(g.bc[bc_marker(g, 1, q.rs)], gamma, 0, u, t);
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = t[v];
        } else if (q.yside != 0) {
            R t[NV];
            bc_image(g.bc[bc_marker(g, q.yside < 0 ? 2 : 3, xs)], gamma, 1, u, t);
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = t[v];
        }
    };
    auto load_row = [&](int r, R (&u)[NV]) {
        RawRow q;
        issue_row(r, q);
        complete_row(q, u);
    };

    // FP32-residual SSP stages with the state kept in double: stage 2 also writes the increment
    // D = U2 - U^n = dt/4 (L(U^n) + L(U1)), and stage 3 forms U^{n+1} = U^n + 2/3 (D + dt L(U2))
    // from it, so U^n is never rounded to float.  The centre value of L(U^n) (stage 2) or of D
    // (stage 3) is loaded with the centre U^n, ahead of the finalize that uses it.
    constexpr bool INC_F2 = MODE == MODE_F2 && sizeof(TOut) == 4, INC_S3 = MODE == MODE_S3 && sizeof(TIn) == 4;
    const bool inc = (INC_F2 && a.dout != nullptr) || (INC_S3 && a.din != nullptr);
    using IncT = typename std::conditional<MODE == MODE_F2, TOut, TIn>::type;
    const IncT* incsrc = (MODE == MODE_F2) ? (const IncT*)a.lin : (const IncT*)a.din;
    auto load_inc = [&](unsigned off, IncT (&d)[NV]) {
        if (inc) {
#pragma unroll
            for (int v = 0; v < NV; v++) d[v] = ld_state<GP>(incsrc, P, v, off);
        }
    };

    // write the result for a finalized cell: the stage update, or L = R/V (MODE_F1)
    auto finish = [&](unsigned off, const R (&u)[NV], const R (&res0)[NV], const double (&un)[NV],
                      const IncT (&dc)[NV], double ivol, double vol) {
        R res[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) res[v] = res0[v];
        if constexpr (VIS) add_sources(a.v, u, vol, res);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if constexpr (MODE == MODE_RN) {
                if (out_lane) rn[v] += vol * __builtin_fabs(double(res[v]));
            } else if constexpr (MODE == MODE_F1) {
                if (out_lane) store(v, off, TOut(double(res[v]) * ivol));
            } else {
                if (inc && MODE == MODE_F2) {
                    const double d = 0.25 * dt * (double(dc[v]) + double(res[v]) * ivol);
                    if (out_lane) {
                        st_state_stream<GP>((TOut*)a.dout, P, v, off, TOut(d));
                        store(v, off, TOut(un[v] + d));
                    }
                } else if (inc && MODE == MODE_S3) {
                    const double d = double(dc[v]);
                    const double q = un[v] + (2.0 / 3.0) * (d + dt * (double(res[v]) * ivol));
                    if (out_lane) store(v, off, TOut(q));
                } else {
                    double q = double(u[v]) + dt * (double(res[v]) * ivol);
                    if constexpr (MT::use_un) q = MT::ca * un[v] + MT::cb * q;
                    if (out_lane) store(v, off, TOut(q));
                }
            }
        }
    };


    if constexpr (DIM == 1) {
        // ---- one row ----
        R c[NV], nb[2][NV], face[2][NV];
        load_row(0, c);
        double un[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) un[v] = MT::use_un ? ld_state<GP>(a.un, P, v, unsigned(xs)) : 0.0;
        IncT dc[NV];
        load_inc(unsigned(xs), dc);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            nb[0][v] = dpp_from_left(c[v]);
            nb[1][v] = dpp_from_right(c[v]);
        }
        reconstruct<1, RECON, LIM, R, FLUX != FX_ROE>(c, nb, geo, face);
        R UL[NV], UR[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            UL[v] = dpp_from_left(face[1][v]);
            UR[v] = face[0][v];
        }
        if (!g.periodic[0]) {
            if (xg == 0) bc_image(g.bc[bc_marker(g, 0, 0)], gamma, 0, UR, UL);
            if (xg == nx) bc_image(g.bc[bc_marker(g, 1, 0)], gamma, 0, UL, UR);
        }
        const R Sx = R(1);
        R lam;
        R F[NV];
        if constexpr (MT::flux) face_flux<FLUX, 0, R>(gamma, eps, UL, UR, F, lam);
        else lam = face_eigen<FLUX, 0, R>(gamma, UL, UR);
        if constexpr (VIS && MT::flux) {
            if (a.v.viscous) {
                const unsigned fi = unsigned(xg < 0 ? 0 : (xg > nx ? nx : xg));
#pragma unroll
                for (int v = 1; v < NV; v++) F[v] -= R(ld_plane(a.v.fv[0], a.v.nface[0], v - 1, fi));
            }
        }
        if constexpr (MT::sr) {
            R lam_r = dpp_from_right(lam);
            double sr;
            {
#pragma clang fp contract(off)
                sr = double(Sx * lam) + double(Sx * lam_r);
            }
            if constexpr (VIS) {
                const double rl = dpp_from_right(double(UL[0])), rr = dpp_from_right(double(UR[0]));
                sr = sr + viscous_sr(rl, rr, g.gamma, a.v.mu, double(Sx), hx);
            }
            if (out_lane) block_min = dt_candidate(hx, sr);
            if (ghost_x_lane) block_min = __builtin_fmin(block_min, dt_candidate(ld_geo_uniform<GP>(g.h[0], nx + 1), double(Sx * lam)));
        }
        if constexpr (MT::flux) {
            R res[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                R Fr = dpp_from_right(F[v]);
                res[v] = -((Fr - F[v]) * Sx);
            }
            finish(unsigned(xs), c, res, un, dc, ihx, hx);
        }
    } else {
        // ---- march along y ----
        const int y0 = yb * a.ty;
        const int yend = (y0 + a.ty < ny) ? y0 + a.ty : ny;
        RawRow Pn;              // prefetched stencil row (completed and entered into the window one
                                // iteration later)
        double Nc[NV];          // U^n of the row finalized in this iteration (SSP stages 2, 3)
        IncT Dc[NV];            // its L(U^n) or increment (FP32-residual stages 2, 3)
        R ytop[NV];             // +y face state of the previously reconstructed row
        // The three-row window (rows rc-1, rc, rc+1) and the values carried to the next row live
        // in LDS, one set per wavefront; no other lane reads them, so no barrier is needed.  The
        // window rotates by slot index instead of register moves, and out of the register file
        // these 20 doubles per lane (40 KB per 4-wavefront block in FP64) leave the occupancy to
        // the register count.
        __shared__ R ring[4][3][NV][64];  // [wavefront][slot][variable][lane]
        // residual carried to the pending row: its x part plus S_y times the y-flux through its
        // bottom face; the finalize subtracts S_y times the flux through its top face
        __shared__ R carry[4][NV][64];
        const int wl = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
        R (&cy)[NV][64] = carry[wl];
        int sm = 0, s0 = 1, sp = 2;  // slots of rows rc-1, rc, rc+1
        // x-neighbours are read from the same LDS rows at lane -+ 1 (the end lanes, halo lanes
        // whose results are discarded, read themselves): LDS reads instead of VALU lane shifts
        const int lm = lane > 0 ? lane - 1 : 0, lp = lane < 63 ? lane + 1 : 63;
        double srp = 0.0;       // spectral radius of the pending row without its top face
        R lam_gb = R(0);
        {
            R t0[NV], t1[NV];
            load_row(y0 - 2, t0);
            load_row(y0 - 1, t1);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                ring[wl][sm][v][lane] = t0[v];
                ring[wl][s0][v][lane] = t1[v];
            }
        }
        issue_row(y0, Pn);
#pragma unroll
        for (int v = 0; v < NV; v++) Nc[v] = 0.0;
        const int nit = yend - y0 + 2;
        for (int t = 0; t < nit; t++) {
            const int rc = y0 - 1 + t;  // row being reconstructed
            {
                R un1[NV];
                complete_row(Pn, un1);
#pragma unroll
                for (int v = 0; v < NV; v++) ring[wl][sp][v][lane] = un1[v];
            }
            if (t + 1 < nit) issue_row(rc + 2, Pn);
            if constexpr (MT::use_un) {
                // U^n of the row finalized below, loaded now: the reconstruction and the y-face
                // flux cover its latency (it was read three iterations ago as a stencil row, so
                // it is normally an L2 hit), and only one buffer is live
                if (t >= 2) {
                    const unsigned offn = unsigned(rc - 1) * unsigned(nx) + unsigned(xs);
#pragma unroll
                    for (int v = 0; v < NV; v++) Nc[v] = ld_state<GP>(a.un, P, v, offn);
                    load_inc(offn, Dc);
                }
            }
            int rcs, rcside;
            map_index(rc, ny, g.periodic[1], rcs, rcside);
            const double hy = ld_geo_uniform<GP>(g.h[1], rcs + 1);
            geo.hh[1] = R(0.5 * hy);
            geo.rc[1] = R(ld_geo_uniform<GP>(g.rc[1], rcs));
            geo.rl[1] = R(ld_geo_uniform<GP>(g.rl[1], rcs));
            geo.rr[1] = R(ld_geo_uniform<GP>(g.rr[1], rcs));
            if (ev_cells) geo.eps2 = R(ev_x * ld_geo_uniform<GP>(g.ev[1], rcs));
            R nb[4][NV], face[4][NV];
            {
                R U0[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    U0[v] = ring[wl][s0][v][lane];
                    nb[0][v] = ring[wl][s0][v][lm];
                    nb[1][v] = ring[wl][s0][v][lp];
                    nb[2][v] = ring[wl][sm][v][lane];
                    nb[3][v] = ring[wl][sp][v][lane];
                }
                reconstruct<2, RECON, LIM, R, FLUX != FX_ROE>(U0, nb, geo, face);
            }
            if (t >= 1) {
                // face between rows rc-1 and rc
                R GL[NV], GR[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) { GL[v] = ytop[v]; GR[v] = face[2][v]; }
                if (!g.periodic[1]) {
                    if (rc == 0) bc_image(g.bc[bc_marker(g, 2, xs)], gamma, 1, GR, GL);
                    if (rc == ny) bc_image(g.bc[bc_marker(g, 3, xs)], gamma, 1, GL, GR);
                }
                const R Sy = R(hx);  // face area of a y-face (ERK:1268)
                R G[NV], lam;
                if constexpr (MT::flux) face_flux<FLUX, 1, R>(gamma, eps, GL, GR, G, lam);
                else lam = face_eigen<FLUX, 1, R>(gamma, GL, GR);
                if constexpr (VIS && MT::flux) {
                    if (a.v.viscous) {
                        const unsigned fi = unsigned(rc) * unsigned(nx) + unsigned(xs);
#pragma unroll
                        for (int v = 1; v < NV; v++) G[v] -= R(ld_plane(a.v.fv[1], a.v.nface[1], v - 1, fi));
                    }
                }
                if constexpr (MT::sr) {
                    if (t >= 2) {
                        double sr;
                        {
#pragma clang fp contract(off)
                            sr = srp + double(Sy * lam);
                        }
                        if constexpr (VIS)
                            sr = sr + viscous_sr(double(GL[0]), double(GR[0]), g.gamma, a.v.mu, double(Sy), hx * g.h[1][rc]);
                        if (out_lane) block_min = __builtin_fmin(block_min, dt_candidate(hx * ld_geo_uniform<GP>(g.h[1], rc), sr));
                    }
                    if (rc == ny && out_lane)  // top ghost row: only this face (ERK:1305)
                        block_min = __builtin_fmin(block_min, dt_candidate(hx * ld_geo_uniform<GP>(g.h[1], ny + 1), double(Sy * lam)));
                    lam_gb = lam;
                }
                if constexpr (MT::flux) {
                    if (t >= 2) {
                        // finalize row rc-1 = the row in slot sm
                        const int rf = rc - 1;
                        const double ivol = ihx * ld_geo_uniform<GP>(g.ih[1], rf + 1);
                        R res[NV], Um[NV];
#pragma unroll
                        for (int v = 0; v < NV; v++) {
                            res[v] = cy[v][lane] - G[v] * Sy;
                            Um[v] = ring[wl][sm][v][lane];
                        }
                        finish(unsigned(rf) * unsigned(nx) + unsigned(xs), Um, res, Nc, Dc, ivol, hx * ld_geo_uniform<GP>(g.h[1], rf + 1));
                    }
#pragma unroll
                    for (int v = 0; v < NV; v++) cy[v][lane] = G[v] * Sy;
                }
                if (rc < yend) {
                    // x-faces of row rc: each lane evaluates its minus face.  The plus-face
                    // states of the left neighbours, then the fluxes and eigenvalues of the right
                    // neighbours, pass through the LDS slot of row rc-1, free after its finalize
                    // (LDS operations of one wavefront complete in order).
                    R (&xch)[NV][64] = ring[wl][sm];
                    R UL[NV], UR[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) xch[v][lane] = face[1][v];
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        UL[v] = xch[v][lm];
                        UR[v] = face[0][v];
                    }
                    if (!g.periodic[0]) {
                        if (xg == 0) bc_image(g.bc[bc_marker(g, 0, rcs)], gamma, 0, UR, UL);
                        if (xg == nx) bc_image(g.bc[bc_marker(g, 1, rcs)], gamma, 0, UL, UR);
                    }
                    const R Sx = R(hy);  // face area of an x-face (ERK:1189)
                    R F[NV], lamx;
                    if constexpr (MT::flux) face_flux<FLUX, 0, R>(gamma, eps, UL, UR, F, lamx);
                    else lamx = face_eigen<FLUX, 0, R>(gamma, UL, UR);
                    if constexpr (VIS && MT::flux) {
                        if (a.v.viscous) {
                            const unsigned fi = unsigned(rcs) * unsigned(nx + 1) + unsigned(xg < 0 ? 0 : (xg > nx ? nx : xg));
#pragma unroll
                            for (int v = 1; v < NV; v++) F[v] -= R(ld_plane(a.v.fv[0], a.v.nface[0], v - 1, fi));
                        }
                    }
                    if constexpr (MT::flux) {
#pragma unroll
                        for (int v = 0; v < NV; v++) xch[v][lane] = F[v];
#pragma unroll
                        for (int v = 0; v < NV; v++) cy[v][lane] = cy[v][lane] - (xch[v][lp] - F[v]) * Sx;
                    }
                    if constexpr (MT::sr) {
                        xch[0][lane] = lamx;
                        const R lam_r = xch[0][lp];
                        {
#pragma clang fp contract(off)
                            srp = double(Sx * lamx) + double(Sx * lam_r);
                        }
                        if constexpr (VIS) {
                            const double rl = dpp_from_right(double(UL[0])), rr = dpp_from_right(double(UR[0]));
                            srp = srp + viscous_sr(rl, rr, g.gamma, a.v.mu, double(Sx), hx * hy);
                        }
                        {
#pragma clang fp contract(off)
                            srp = srp + double(Sy * lam_gb);
                        }
                        if (ghost_x_lane)  // right ghost column (ERK:1225)
                            block_min = __builtin_fmin(block_min, dt_candidate(ld_geo_uniform<GP>(g.h[0], nx + 1) * hy, double(Sx * lamx)));
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < NV; v++) ytop[v] = face[3][v];
            const int st = sm;
            sm = s0;
            s0 = sp;
            sp = st;
        }
    }

}

// ---- single-solve kernel: one wavefront per tile, block-level dt reduction ----------------------
template <int DIM, int RECON, int LIM, int FLUX, int MODE, typename R, typename TIn, typename TOut, int VIS = 0>
__global__ void __launch_bounds__(256) fvm_stage(KArgs a) {
    static_assert(DIM == 1 || DIM == 2, "3-D uses its own kernel");
    using MT = ModeTraits<MODE>;
    StepCtl* ctl = a.ctl;
    if constexpr (MT::sr) {
        if (!ctl->active) {
            if (blockIdx.x == 0 && threadIdx.x == 0) ctl->go = 0;
            return;
        }
    } else {
        if (!ctl->go) return;
        if constexpr (MODE == MODE_RN) {
            if (ctl->active) return;  // not the last step of this call
        }
    }
    const int lane = threadIdx.x & 63;
    // the wavefront index is the same in all its lanes: readfirstlane lets the compiler keep the
    // tile, its rows and every row's geometry in scalar registers (scalar loads, whose waits do
    // not also wait for the vector loads of the stencil prefetch)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    const int tile = xcd_remap(blockIdx.x, gridDim.x) * nwaves + wave;
    if constexpr (MODE != MODE_RN) {
        if (tile >= a.ntiles) return;  // only the norms reduce across the block
    }
    double block_min = __builtin_inf();
    double rn[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    stage_tile<DIM, RECON, LIM, FLUX, MODE, R, TIn, TOut, VIS>(a, tile < a.ntiles ? tile : -1, lane, block_min, rn);
    if constexpr (MT::sr) {
        // store the wavefront's minimum; dt_finalize_kernel turns the tiles' minima into dt after
        // this launch
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) block_min = __builtin_fmin(block_min, __shfl_xor(block_min, o));
        if (lane == 0) st_plane(a.dt_part, 0, 0, unsigned(tile), block_min);
    }
    if constexpr (MODE == MODE_RN) {
        __shared__ double rred[4 * NV];
        rn_reduce_finalize(ctl, a.rn_part, rn, rred, lane, wave, nwaves);
    }
}

// ---- small 1-D forward-Euler solves: a chunk of steps in one workgroup -------------------------
// A grid of a few strips is launch-bound (two short kernels per step).  One workgroup of W
// wavefronts runs up to `cap` steps of a ts_step call instead: the dt pass over its strips (each
// wavefront takes strips w, w + W, ...), the workgroup minimum in LDS, the finalize by one
// thread, a barrier, the forward-Euler stage into the other state buffer and a barrier before
// the next dt pass reads it.  The two directions of the state ping-pong come as two kernel
// arguments (a0: buffer 0 -> 1, a1: 1 -> 0), so that nothing is copied per step.  Every
// wavefront leaves the loop together: after a step that ends the run (stop rule, error) or
// after `cap` steps.  The control block lives in LDS for the chunk (thread 0 copies it in at the
// start and back at the end), so the per-step bookkeeping -- the stop test, the dt update and
// the stage's read of dt -- makes no global round trip.
// W = 16 covers the 17 strips of 1000 cells in two rounds instead of three; the exact solver's
// registers do not fit the 128 per lane of a 1024-thread workgroup, so Godunov keeps W = 8.
// Grids of at most WG1D_LDS_CELLS cells keep both state buffers in LDS for the chunk (copied in
// from the input buffer at the start, the final state copied out to the buffer the host expects
// it in at the end): the dt pass and the stage then make no global round trip at all.
template <int FLUX>
constexpr int wg1d_waves = FLUX == FX_GODUNOV ? 8 : 16;
constexpr int WG1D_LDS_CELLS = 1984;  // 2 buffers x 5 planes x 1984 doubles = 158720 bytes

// one step; false when the loop has to end
template <int RECON, int LIM, int FLUX, bool LDS>
__device__ __forceinline__ bool euler1d_step(const KArgs& a, StepCtl* ctl, double* red, int* run, int lane,
                                             int wave, const double* sin, double* sout) {
    if (threadIdx.x == 0) *run = ctl->active;
    __syncthreads();
    if (!*run) {
        if (threadIdx.x == 0) ctl->go = 0;
        return false;
    }
    double m = __builtin_inf();
    for (int tile = wave; tile < a.ntiles; tile += wg1d_waves<FLUX>) {
        double mt = __builtin_inf();
        stage_tile<1, RECON, LIM, FLUX, MODE_DT, double, double, double, 0, false, LDS>(
            a, tile, lane, mt, nullptr, 0.0, sin, sout, WG1D_LDS_CELLS);
        m = __builtin_fmin(m, mt);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = __builtin_fmin(m, __shfl_xor(m, o));
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double bmin = red[0];
        for (int w = 1; w < wg1d_waves<FLUX>; w++) bmin = __builtin_fmin(bmin, red[w]);
        // the whole grid is this workgroup: no atomics on the control block
        dt_apply(ctl, a.g.cfl, bmin);
        *run = ctl->go;
    }
    __syncthreads();
    if (!*run) return false;  // non-physical step: the state of the last good step stays in place
    const double dt = ctl->dt;
    for (int tile = wave; tile < a.ntiles; tile += wg1d_waves<FLUX>) {
        double unused = 0.0;
        stage_tile<1, RECON, LIM, FLUX, MODE_S1, double, double, double, 0, true, LDS>(
            a, tile, lane, unused, nullptr, dt, sin, sout, WG1D_LDS_CELLS);
    }
    __syncthreads();
    return true;
}

template <int RECON, int LIM, int FLUX, bool LDS>
__global__ void __launch_bounds__(64 * wg1d_waves<FLUX>) fvm_euler1d_workgroup(KArgs a0, KArgs a1, int cap) {
    constexpr int NT = 64 * wg1d_waves<FLUX>;
    __shared__ double red[wg1d_waves<FLUX>];
    __shared__ int run;
    __shared__ StepCtl ctl;  // a0.ctl and a1.ctl are the same block
    __shared__ double st[LDS ? 2 * NV * WG1D_LDS_CELLS : 1];
    double* b0 = st;
    double* b1 = st + (LDS ? NV * WG1D_LDS_CELLS : 0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ncell = a0.g.n[0];  // at most WG1D_LDS_CELLS when LDS (checked by the launcher)
    if (threadIdx.x == 0) ctl = *a0.ctl;  // read only by thread 0 until the first barrier
    if constexpr (LDS) {
        const double* g = (const double*)a0.uin;
        for (int c = threadIdx.x; c < ncell; c += NT)
#pragma unroll
            for (int v = 0; v < NV; v++) b0[v * WG1D_LDS_CELLS + c] = g[(size_t)v * a0.plane + c];
        __syncthreads();
    }
    int done = 0;  // uniform: every wavefront takes the same branches
    for (int it = 0; it < cap; it += 2) {
        if (!euler1d_step<RECON, LIM, FLUX, LDS>(a0, &ctl, red, &run, lane, wave, b0, b1)) break;
        done++;
        if (it + 1 >= cap) break;
        if (!euler1d_step<RECON, LIM, FLUX, LDS>(a1, &ctl, red, &run, lane, wave, b1, b0)) break;
        done++;
    }
    if constexpr (LDS) {
        // the state after `done` steps is in b0 (even) or b1 (odd); the host expects it in the
        // input buffer after an even number of steps and in the other one after an odd number.
        // Every step ends at a barrier, and one that stops early writes no state after it.
        const double* src = (done & 1) ? b1 : b0;
        double* g = (done & 1) ? (double*)a0.uout : (double*)const_cast<void*>(a0.uin);
        for (int c = threadIdx.x; c < ncell; c += NT)
#pragma unroll
            for (int v = 0; v < NV; v++) g[(size_t)v * a0.plane + c] = src[v * WG1D_LDS_CELLS + c];
    }
    if (threadIdx.x == 0) *a0.ctl = ctl;  // the other threads only ever read the LDS copy
}

// ---- batched kernel: a persistent grid over a table of (problem, tile) pairs -----------------
// Every problem of the batch keeps its own KArgs (state, geometry, control block).  Wavefront w of
// the persistent grid takes tiles w, w + W, w + 2W, ... of the table (W = wavefronts in the grid):
// no claim atomics, which serialise at one address at tens of ns each, and every wavefront's
// loop ends after at most ceil(ntiles / W) tiles.  Neighbouring wavefront ids sit on one XCD, so
// neighbouring tiles of a problem share an L2.  Stage 1 stores each table entry's wavefront-level
// minimum of V/SR in `part` (by entry), and dt_finalize_batched turns each problem's entries into
// its dt.  A cell's arithmetic depends only on its own problem, and the minimum is exact, so a
// result is the same alone or in a mix.

template <int DIM, int RECON, int LIM, int FLUX, int MODE, typename R, typename TIn, typename TOut>
__global__ void __launch_bounds__(256) fvm_stage_batched(const KArgs* __restrict__ descs,
                                                         const int2* __restrict__ table,
                                                         const int* __restrict__ ntiles_p,
                                                         double* __restrict__ part) {
    using MT = ModeTraits<MODE>;
    const int lane = threadIdx.x & 63;
    // loop control in SGPRs: every quantity below is uniform over the wavefront
    const int ntiles = __builtin_amdgcn_readfirstlane(*ntiles_p);
    const int nwpb = blockDim.x >> 6;
    const int stride = gridDim.x * nwpb;
    const int w0 = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * nwpb + (threadIdx.x >> 6));
    for (int t = w0; t < ntiles; t += stride) {
        const int2 e = table[t];
        const KArgs& a = descs[e.x];
        StepCtl* ctl = a.ctl;
        if constexpr (MT::sr) {
            if (!ctl->active) {
                if (e.y == 0 && lane == 0) ctl->go = 0;
                continue;
            }
        } else {
            if (!ctl->go) continue;
        }
        double m = __builtin_inf();
        if constexpr (DIM == 1) {
            // a 1-D work item is a group of a.ty consecutive strips, walked by one wavefront
            const int s0 = e.y * a.ty;
            const int s1 = (s0 + a.ty < a.ntiles) ? s0 + a.ty : a.ntiles;
            for (int sidx = s0; sidx < s1; sidx++) {
                double ms = __builtin_inf();
                stage_tile<DIM, RECON, LIM, FLUX, MODE, R, TIn, TOut, 0, false, false, true>(a, sidx, lane, ms);
                m = __builtin_fmin(m, ms);
            }
        } else {
            stage_tile<DIM, RECON, LIM, FLUX, MODE, R, TIn, TOut, 0, false, false, true>(a, e.y, lane, m);
        }
        if constexpr (MT::sr) {
            // the entry's minimum; dt_finalize_batched reduces each solve's entries after the launch
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) m = __builtin_fmin(m, __shfl_xor(m, o));
            if (lane == 0) part[t] = m;
        }
    }
}

}  // namespace tsd
