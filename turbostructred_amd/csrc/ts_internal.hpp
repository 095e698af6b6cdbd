// ts_internal.hpp -- host-side state of a solver context (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ts_fvm.h"
#include "fvm_dispatch.hpp"

struct ts_ctx {
    ts_config cfg;
    std::string err;
    int nd = 1;
    int n[3] = {1, 1, 1};
    size_t ncell = 0;
    hipStream_t stream = nullptr;

    // device state: dU[0] is U^n (double); dU[1], dU[2] are stage arrays (double, or float in
    // the FP32-residual SSP form, which also keeps the stage-2 increment U2 - U^n in dU[3]).
    // Forward Euler ping-pongs between dU[0] and dU[1]; cur is the buffer holding the current state.
    void* dU[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t stage_elem = 8;
    int cur = 0;
    // SSP-RK3 with the dt pass fused into stage 1 (dU[1] then holds L(U^n)); off when the
    // environment sets TS_SEPARATE_DT=1
    bool fused_dt = true;
    // small 1-D forward-Euler solves run chunks of steps in one workgroup (no launch per step);
    // also off with TS_SEPARATE_DT=1
    bool euler1d_wg = false;

    // geometry on the device
    std::vector<double*> dbufs;  // everything to free
    tsd::GridArgs grid{};
    std::vector<double> X[3], H[3];
    std::vector<uint8_t> hmarker[6];  // host copies of per-face markers (ghost images on download)

    tsd::StepCtl* dctl = nullptr;
    tsd::StepCtl* hctl = nullptr;  // pinned mirror
    double* dt_log = nullptr;
    long long dt_log_cap = 0, dt_log_base = 0;

    tsd::KernelSet ks;
    int ty = 16;
    int nstrips = 1, ntiles = 1;
    dim3 grid_dim, block_dim;

    // host view of the step bookkeeping
    double h_time = 0, h_dt = 0;
    long long h_iter = 0;
    int h_steps_done = 0;

    // kernel timing
    bool timing = false;
    // per kind: dt, stages 1-3, (unused), residual norms, viscous pre-pass
    double t_ms[7] = {0, 0, 0, 0, 0, 0, 0};
    long long t_n[7] = {0, 0, 0, 0, 0, 0, 0};
    // viscous terms / body forces / viscous spectral radius: the VIS kernel instances run
    bool vis = false;
    tsd::VisArgs visargs{};
    unsigned* d_cells = nullptr;  // ts_get_cells: cell indices and gathered values
    double* d_cell_vals = nullptr;
    size_t cells_cap = 0;
    bool norms = false;         // residual-norm pass after the first kernel of every step
    double* rn_part = nullptr;  // its per-block partial sums
    double* dt_part = nullptr;  // the dt passes' per-tile minima of V/SR
    double h_rnorm[5] = {0, 0, 0, 0, 0};
    std::vector<hipEvent_t> events;

    // hipGraph of a chunk of steps (built lazily per (chunk length, start buffer))
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    bool use_graphs = true;

    double* d_aos = nullptr;  // device copy of the reference AoS state (uploads / downloads)
    size_t device_bytes = 0;

    // Borrowed-memory mode (ts_step_noalloc): device buffers are carved out of `arena` (a fixed
    // region the library owns for the process), the state buffer dU[0] and the stream are the
    // caller's, and the control block's host mirror is hctl_local.  Nothing is allocated or freed.
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    tsd::StepCtl hctl_local{};
};

namespace tsr {
std::string hip_msg(hipError_t e, const char* what);
void build_axis(const ts_config& c, int a, std::vector<double>& X, std::vector<double>& H);
bool uniform_axis(const std::vector<double>& H);
int recon_variant(const ts_config& c);
// the configuration checks of ts_create: TS_OK, or an error code with its message in err
int validate(const ts_config* c, std::string& err);
// Venkatakrishnan eps^2 = (K (hx hy hz)^(1/nd))^3, K = 5 (Limiters.h:93-96, Constants.h:11)
double venkat_eps2(int nd, double hx, double hy, double hz);
// per-axis factors (K h_i)^(3/nd) of that value for the n interior cells of an axis whose sizes
// start at h (null for an inactive axis: one factor 1)
void venkat_axis_factors(int nd, int n, const double* h, std::vector<double>& out);
bool lookup_kernels(int dim, int flux, int rv, int precision, tsd::KernelSet* ks);
void host_bc_image(const ts_bc_config& bc, double g, int axis, const double* u, double* o);
// output rows per tile; `slots` = tiles the GPU holds at once (0: unknown)
int tuned_rows_per_tile(const ts_config& c, long long slots = 0);
// the calling thread's error text for entry points without a context (ts_last_error(NULL))
void set_thread_error(const std::string& m);
// the reference AoS layout (5 values per cell, ghost layers on active axes) <-> SoA interior
// planes, on the device; the AoS ghost entries are neither read nor written
size_t aos_doubles(int nd, const int n[3]);
hipError_t device_aos_to_soa(const double* d_aos, double* d_soa, int nd, const int n[3], hipStream_t stream);
hipError_t device_soa_to_aos(const double* d_soa, double* d_aos, int nd, const int n[3], hipStream_t stream);
}  // namespace tsr

size_t ts_state_size_cfg(const ts_config* c);
int init_device_ctx(ts_ctx* c, const ts_config* cfg);
void destroy_device_ctx(ts_ctx* c);
void fill_ghost_images(const ts_ctx* c, double* aos);
int get_cells(ts_ctx* c, const int* ijk, size_t n, double* out);
int run_steps(ts_ctx* c, int max_steps, double t_max, double next_snap);
