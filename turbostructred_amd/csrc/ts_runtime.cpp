// ts_runtime.cpp -- host side of the C ABI (include/ts_fvm.h) for one solve per context.
//
// Replaces Kernel::Init / ExplicitRungeKuttaFVM::InitializeMethod (kernel.h:178-281,
// ERK:39-75), SetInitialConditions' upload, IterationStep (ERK:1440-1491) and the termination
// rule of Kernel::Run (kernel.h:1834-1846).  State lives on the device as five SoA planes of
// interior cells; the reference's AoS array with ghost layers only exists at the boundary
// (ts_set_state / ts_get_state).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ts_fvm.h"
#include "fvm_dispatch.hpp"
#include "ts_internal.hpp"

using namespace tsd;

namespace tsr {

std::string hip_msg(hipError_t e, const char* what) {
    return std::string(what) + ": " + hipGetErrorString(e);
}

// Build per-axis geometry exactly as grid.h:124-218 does for a uniform block (running sums of
// the step), or take the caller's arrays.
void build_axis(const ts_config& c, int a, std::vector<double>& X, std::vector<double>& H) {
    int n = a == 0 ? c.nx : (a == 1 ? c.ny : c.nz);
    double L = a == 0 ? c.lx : (a == 1 ? c.ly : c.lz);
    int g = c.ghost;
    X.assign(n + 2 * g, 0.0);
    H.assign(n + 2 * g, 0.0);
    if (c.coord[a] && c.h[a]) {
        for (int i = 0; i < n + 2 * g; i++) {
            X[i] = c.coord[a][i];
            H[i] = c.h[a][i];
        }
        return;
    }
    double stp = L / n;
    double xc = -(g * stp) + 0.5 * stp;
    int k = 0;
    for (; k < g; k++) { X[k] = xc; H[k] = stp; xc += stp; }
    for (int i = 0; i < n; i++) { X[k + i] = xc; H[k + i] = stp; xc += 0.5 * stp * (1.0 + 1.0); }
    k += n;
    for (int i = 0; i < g; i++) { X[k + i] = xc; H[k + i] = stp; xc += stp; }
}

bool uniform_axis(const std::vector<double>& H) {
    for (double h : H)
        if (h != H[0]) return false;
    return true;
}

int validate(const ts_config* c, std::string& err) {
    if (!c) { err = "null config"; return TS_ERR_INVALID; }
    if (c->ndims < 1 || c->ndims > 3) { err = "ndims must be 1, 2 or 3"; return TS_ERR_INVALID; }
    if (c->ghost != 1) { err = "DummyLayerSize (ghost) must be 1"; return TS_ERR_UNSUPPORTED; }
    int n[3] = {c->nx, c->ny, c->nz};
    for (int a = 0; a < c->ndims; a++)
        if (n[a] < 2) { err = "every active axis needs at least 2 cells"; return TS_ERR_INVALID; }
    for (int a = c->ndims; a < 3; a++)
        if (n[a] != 1) { err = "inactive axes must have 1 cell"; return TS_ERR_INVALID; }
    // a state plane (one variable, one double per cell) is addressed with 32-bit byte offsets
    if ((long long)n[0] * n[1] * n[2] * 8 >= (1ll << 32)) {
        err = "more than 2^29 cells: a state plane would exceed 4 GiB";
        return TS_ERR_UNSUPPORTED;
    }
    if (!(c->gamma > 1.0)) { err = "gamma must exceed 1"; return TS_ERR_INVALID; }
    if (!(c->cfl > 0.0)) { err = "CFL must be positive"; return TS_ERR_INVALID; }
    if (c->riemann != TS_RIEMANN_ROE && c->riemann != TS_RIEMANN_GODUNOV && c->riemann != TS_RIEMANN_HLLC) {
        err = "unknown Riemann solver";
        return TS_ERR_INVALID;
    }
    if (c->recon < 0 || c->recon > 3) { err = "unknown reconstruction"; return TS_ERR_INVALID; }
    if (c->limiter < 0 || c->limiter > 1) { err = "unknown limiter"; return TS_ERR_INVALID; }
    if (c->rk_stages != 1 && c->rk_stages != 3) { err = "rk_stages must be 1 or 3"; return TS_ERR_INVALID; }
    if (c->precision != TS_FP64 && c->precision != TS_FP32_RESIDUAL) { err = "unknown precision"; return TS_ERR_INVALID; }
    if (c->n_bc < 0 || c->n_bc > TS_MAX_BC) { err = "n_bc out of range"; return TS_ERR_INVALID; }
    for (int m = 0; m < c->n_bc; m++)
        if (c->bc[m].type < TS_BC_WALL || c->bc[m].type > TS_BC_NATURAL) { err = "unknown BC type"; return TS_ERR_INVALID; }
    if (c->viscous || c->external_force || c->external_accel || c->viscosity != 0.0) {
        if (!(c->viscosity >= 0.0) || !(c->viscosity < 1e300)) { err = "viscosity must be finite and >= 0"; return TS_ERR_INVALID; }
        for (int a = 0; a < 3; a++)
            if (!(c->sigma[a] - c->sigma[a] == 0.0) || !(c->accel[a] - c->accel[a] == 0.0)) {
                err = "Sigma and UniformAcceleration must be finite";
                return TS_ERR_INVALID;
            }
        if (c->precision != TS_FP64) { err = "viscous terms and body forces run in FP64 only"; return TS_ERR_UNSUPPORTED; }
    }
    for (int a = 0; a < c->ndims; a++) {
        if (c->periodic[a]) continue;
        for (int s = 0; s < 2; s++) {
            int side = 2 * a + s;
            if (c->face_marker[side]) continue;
            if (c->side_marker[side] < 0 || c->side_marker[side] >= c->n_bc) {
                err = "side marker refers to a missing boundary condition";
                return TS_ERR_INVALID;
            }
        }
    }
    return TS_OK;
}

double venkat_eps2(int nd, double hx, double hy, double hz) {
    return std::pow(5.0 * std::pow(hx * hy * hz, 1.0 / nd), 3.0);
}

void venkat_axis_factors(int nd, int n, const double* h, std::vector<double>& out) {
    if (!h) {
        out.assign(1, 1.0);
        return;
    }
    out.resize(n);
    const double e = 3.0 / nd;
    const double kf = std::pow(5.0, e);
    for (int i = 0; i < n; i++) out[i] = kf * std::pow(h[i], e);
}

int recon_variant(const ts_config& c) {
    if (c.recon == TS_RECON_LINEAR_LIM) return c.limiter == TS_LIMITER_VENKAT ? 4 : 3;
    return c.recon;
}

bool lookup_kernels(int dim, int flux, int rv, int precision, KernelSet* ks) {
    if (dim == 1) {
        if (flux == 0) return get_kernels_d1_f0(rv, precision, ks);
        if (flux == 1) return get_kernels_d1_f1(rv, precision, ks);
        if (flux == 3) return get_kernels_d1_f3(rv, precision, ks);
    } else if (dim == 2) {
        if (flux == 0) return get_kernels_d2_f0(rv, precision, ks);
        if (flux == 1) return get_kernels_d2_f1(rv, precision, ks);
        if (flux == 3) return get_kernels_d2_f3(rv, precision, ks);
    } else {
        if (flux == 0) return get_kernels_d3_f0(rv, precision, ks);
        if (flux == 1) return get_kernels_d3_f1(rv, precision, ks);
        if (flux == 3) return get_kernels_d3_f3(rv, precision, ks);
    }
    return false;
}

// Host copy of the BC images used for the ghost layers of ts_get_state.
void host_bc_image(const ts_bc_config& bc, double g, int axis, const double* u, double* o) {
    for (int v = 0; v < 5; v++) o[v] = u[v];
    switch (bc.type) {
    case TS_BC_SYMMETRY: o[1 + axis] = -u[1 + axis]; break;
    case TS_BC_WALL: o[1] = -u[1]; o[2] = -u[2]; o[3] = -u[3]; break;
    case TS_BC_MOVING_WALL: {
        double r = u[0], vx = u[1] / r, vy = u[2] / r, vz = u[3] / r;
        double re = u[4] - 0.5 * r * (vx * vx + vy * vy + vz * vz);
        double dx = 2 * bc.velocity[0] - vx, dy = 2 * bc.velocity[1] - vy, dz = 2 * bc.velocity[2] - vz;
        o[1] = r * dx; o[2] = r * dy; o[3] = r * dz;
        o[4] = re + 0.5 * r * (dx * dx + dy * dy + dz * dz);
        break;
    }
    case TS_BC_SUBSONIC_INLET: {
        const double* V = bc.velocity;
        o[0] = bc.density; o[1] = bc.density * V[0]; o[2] = bc.density * V[1]; o[3] = bc.density * V[2];
        o[4] = bc.pstatic / (g - 1) + 0.5 * bc.density * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
        break;
    }
    default: break;
    }
}

}  // namespace tsr

using namespace tsr;

size_t ts_state_size_cfg(const ts_config* c) {
    size_t nxa = c->nx + 2, nya = c->ndims > 1 ? c->ny + 2 : 1, nza = c->ndims > 2 ? c->nz + 2 : 1;
    return nxa * nya * nza * 5;
}

#define HIPCHK(ctx, call)                                        \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) {                                  \
            (ctx)->err = hip_msg(e_, #call);                     \
            return TS_ERR_HIP;                                   \
        }                                                        \
    } while (0)

static int ctx_fail(ts_ctx* c, int code, const std::string& m) {
    c->err = m;
    return code;
}

// Message of the calling thread's last failed ts_create, where no context exists to hold it;
// ts_last_error(NULL) returns it.
static thread_local std::string t_create_err;

void tsr::set_thread_error(const std::string& m) { t_create_err = m; }

extern "C" {

int ts_create(const ts_config* cfg, ts_ctx** out) {
    if (!out) {
        t_create_err = "ts_create: null output pointer";
        return TS_ERR_INVALID;
    }
    *out = nullptr;
    std::string err;
    int rc = validate(cfg, err);
    if (rc) {
        t_create_err = "ts_create: " + err;
        return rc;
    }
    ts_ctx* c = new ts_ctx();
    c->cfg = *cfg;
    for (int s = 0; s < 6; s++) c->cfg.face_marker[s] = nullptr;
    for (int a = 0; a < 3; a++) { c->cfg.coord[a] = nullptr; c->cfg.h[a] = nullptr; }
    c->nd = cfg->ndims;
    c->n[0] = cfg->nx;
    c->n[1] = cfg->ny;
    c->n[2] = cfg->nz;
    c->ncell = (size_t)c->n[0] * c->n[1] * c->n[2];
    int r = init_device_ctx(c, cfg);
    if (r != TS_OK) {
        t_create_err = "ts_create: " + c->err;
        destroy_device_ctx(c);
        delete c;
        return r;
    }
    *out = c;
    t_create_err.clear();
    return TS_OK;
}

int ts_destroy(ts_ctx* c) {
    if (!c) return TS_ERR_INVALID;
    destroy_device_ctx(c);
    delete c;
    return TS_OK;
}

const char* ts_last_error(const ts_ctx* c) {
    if (c) return c->err.c_str();
    return t_create_err.empty() ? "null context" : t_create_err.c_str();
}

size_t ts_state_size(const ts_ctx* c) {
    if (!c) return 0;
    size_t nxa = c->n[0] + 2, nya = c->nd > 1 ? c->n[1] + 2 : 1, nza = c->nd > 2 ? c->n[2] + 2 : 1;
    return nxa * nya * nza * 5;
}

// The caller's AoS array crosses PCIe as it is; the layout conversion runs on the device, in a
// device copy of the AoS array allocated on first use.
static int ensure_aos(ts_ctx* c) {
    if (c->d_aos) return TS_OK;
    const size_t bytes = ts_state_size(c) * sizeof(double);
    hipError_t e = hipMalloc((void**)&c->d_aos, bytes);
    if (e != hipSuccess) {
        c->d_aos = nullptr;
        return ctx_fail(c, e == hipErrorOutOfMemory ? TS_ERR_NOMEM : TS_ERR_HIP, tsr::hip_msg(e, "hipMalloc"));
    }
    c->device_bytes += bytes;
    return TS_OK;
}

int ts_set_state(ts_ctx* c, const double* aos, size_t n) {
    if (!c || !aos) return TS_ERR_INVALID;
    if (n != ts_state_size(c)) return ctx_fail(c, TS_ERR_INVALID, "state size mismatch");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int r = ensure_aos(c);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_aos, aos, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, tsr::device_aos_to_soa(c->d_aos, (double*)c->dU[0], c->nd, c->n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cur = 0;
    return TS_OK;
}

int ts_get_state(ts_ctx* c, double* aos, size_t n) {
    if (!c || !aos) return TS_ERR_INVALID;
    if (n != ts_state_size(c)) return ctx_fail(c, TS_ERR_INVALID, "state size mismatch");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int r = ensure_aos(c);
    if (r) return r;
    // interior cells converted on the device into a zeroed AoS array; the host then writes the
    // ghost layers' images
    HIPCHK(c, hipMemsetAsync(c->d_aos, 0, n * sizeof(double), c->stream));
    HIPCHK(c, tsr::device_soa_to_aos((const double*)c->dU[c->cur], c->d_aos, c->nd, c->n, c->stream));
    HIPCHK(c, hipMemcpyAsync(aos, c->d_aos, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    fill_ghost_images(c, aos);
    return TS_OK;
}

int ts_get_cells(ts_ctx* c, const int* ijk, size_t n, double* out) {
    if (!c || (n && (!ijk || !out))) return TS_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return get_cells(c, ijk, n, out);
}

int ts_set_time(ts_ctx* c, double time, long long iteration) {
    if (!c) return TS_ERR_INVALID;
    c->h_time = time;
    c->h_iter = iteration;
    return TS_OK;
}

int ts_enable_residual_norms(ts_ctx* c, int enable) {
    if (!c) return TS_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (enable && !c->rn_part) {
        HIPCHK(c, hipMalloc((void**)&c->rn_part, (size_t)c->grid_dim.x * 5 * sizeof(double)));
        c->device_bytes += (size_t)c->grid_dim.x * 5 * sizeof(double);
    }
    if (c->norms != (enable != 0)) {
        // the captured step graphs hold the launch sequence: re-capture with or without the pass
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < 2; i++) {
            if (c->graph_exec[i]) HIPCHK(c, hipGraphExecDestroy(c->graph_exec[i]));
            c->graph_exec[i] = nullptr;
            c->graph_steps[i] = 0;
        }
    }
    c->norms = enable != 0;
    for (int v = 0; v < 5; v++) c->h_rnorm[v] = 0.0;
    return TS_OK;
}

int ts_enable_dt_log(ts_ctx* c, long long capacity) {
    if (!c || capacity < 0) return TS_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->dt_log) { HIPCHK(c, hipFree(c->dt_log)); c->dt_log = nullptr; }
    c->dt_log_cap = capacity;
    c->dt_log_base = c->h_iter;
    if (capacity > 0) HIPCHK(c, hipMalloc(&c->dt_log, capacity * sizeof(double)));
    return TS_OK;
}

int ts_get_dt_log(ts_ctx* c, double* out, long long capacity, long long* n_out) {
    if (!c || !out || !n_out) return TS_ERR_INVALID;
    long long n = std::min<long long>(c->h_iter - c->dt_log_base, c->dt_log_cap);
    n = std::min(n, capacity);
    if (n < 0) n = 0;
    if (n > 0) {
        HIPCHK(c, hipMemcpy(out, c->dt_log, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    *n_out = n;
    return TS_OK;
}

int ts_enable_kernel_timing(ts_ctx* c, int enable) {
    if (!c) return TS_ERR_INVALID;
    c->timing = enable != 0;
    for (int k = 0; k < 7; k++) { c->t_ms[k] = 0; c->t_n[k] = 0; }
    return TS_OK;
}

int ts_get_kernel_timing(ts_ctx* c, int kind, double* total_ms, long long* launches) {
    if (!c || kind < 0 || kind > 6 || !total_ms || !launches) return TS_ERR_INVALID;
    if (kind == 4) {
        double s = 0;
        long long n = 0;
        for (int k = 0; k < 4; k++) { s += c->t_ms[k]; n += c->t_n[k]; }
        *total_ms = s;
        *launches = n;
    } else {
        *total_ms = c->t_ms[kind];
        *launches = c->t_n[kind];
    }
    return TS_OK;
}

int ts_memory_footprint(const ts_ctx* c, size_t* bytes) {
    if (!c || !bytes) return TS_ERR_INVALID;
    *bytes = c->device_bytes;
    return TS_OK;
}

int ts_algorithmic_bytes_per_cell(const ts_ctx* c, int kind, double* bytes) {
    if (!c || !bytes || kind < 0 || kind > 6 || kind == 5) return TS_ERR_INVALID;
    // one state array = 5 * sizeof(real) per cell; halo re-reads are not counted
    const double d = 40.0;
    const double s = c->cfg.precision == TS_FP32_RESIDUAL ? 20.0 : 40.0;
    double b[5];
    b[0] = d;  // dt pass reads U^n
    if (c->cfg.rk_stages == 1) {
        b[1] = 2 * d;  // read U^n, write U^{n+1}
        b[2] = b[3] = 0;
    } else if (c->fused_dt) {
        b[0] = 0;            // no separate dt pass
        b[1] = d + s;        // read U^n, write L(U^n)
        b[2] = d + s + s;    // read U^n and L(U^n), write U^2
        b[3] = s + d + d;    // read U^2, U^n, write U^{n+1}
        if (c->stage_elem == 4) {
            b[2] += s;  // FP32-residual: also write the increment U^2 - U^n
            b[3] += s;  // and read it back for the double accumulation of U^{n+1}
        }
    } else {
        b[1] = d + s;        // read U^n, write U^1
        b[2] = s + d + s;    // read U^1, U^n, write U^2
        b[3] = s + d + d;    // read U^2, U^n, write U^{n+1}
    }
    // viscous fluxes: each stage also reads 4 face values per axis; the pre-pass writes 3 nd
    // gradients per cell (reading the state), then reads state and gradients and writes the faces
    double pre = 0.0;
    if (c->vis && c->cfg.viscous) {
        const double nd = c->nd;
        for (int k = 1; k <= 3; k++)
            if (b[k] > 0) b[k] += 32.0 * nd;
        pre = (d + 24.0 * nd) + (d + 24.0 * nd + 32.0 * nd);
    }
    b[4] = b[0] + b[1] + b[2] + b[3] + pre * (c->cfg.rk_stages == 1 ? 1 : 3);
    *bytes = kind == 6 ? pre : b[kind];
    return TS_OK;
}

int ts_step(ts_ctx* c, int max_steps, double t_max, double next_snapshot_time, ts_step_info* out) {
    if (!c || max_steps < 0) return TS_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc = run_steps(c, max_steps, t_max, next_snapshot_time);
    if (out) {
        out->time = c->h_time;
        out->dt = c->h_dt;
        out->iteration = c->h_iter;
        for (int v = 0; v < 5; v++) out->residual_norm[v] = c->h_rnorm[v];
        out->steps_done = c->h_steps_done;
    }
    return rc;
}

}  // extern "C"
