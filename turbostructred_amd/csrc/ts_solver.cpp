// ts_solver.cpp -- device set-up, state layout conversion and the step loop of one context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ts_internal.hpp"

using namespace tsd;

#define CHK(call)                                        \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) {                          \
            c->err = tsr::hip_msg(e_, #call);            \
            return TS_ERR_HIP;                           \
        }                                                \
    } while (0)

namespace {

template <typename T>
int dalloc(ts_ctx* c, T** p, size_t count) {
    if (c->arena) {
        // borrowed-memory mode: bump allocation in the library's fixed region, 256-byte aligned
        const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (c->arena_used + bytes > c->arena_cap) {
            c->err = "the problem needs more than the " + std::to_string(c->arena_cap) +
                     " bytes of the allocation-free scratch region";
            return TS_ERR_INVALID;
        }
        *p = (T*)(c->arena + c->arena_used);
        c->arena_used += bytes;
        c->device_bytes += bytes;
        return TS_OK;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        c->err = tsr::hip_msg(e, "hipMalloc");
        return e == hipErrorOutOfMemory ? TS_ERR_NOMEM : TS_ERR_HIP;
    }
    c->dbufs.push_back((double*)q);
    c->device_bytes += count * sizeof(T);
    *p = (T*)q;
    return TS_OK;
}

template <typename T>
int upload(ts_ctx* c, const T* host, size_t count, const T** dev) {
    T* d = nullptr;
    int r = dalloc(c, &d, count);
    if (r) return r;
    CHK(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dev = d;
    return TS_OK;
}

inline size_t gidx(const ts_ctx* c, int i, int j, int k) {
    size_t nxa = c->n[0] + 2, nya = c->nd > 1 ? c->n[1] + 2 : 1;
    return ((size_t)k * nya + j) * nxa + i;
}

constexpr int CHUNK = 256;  // steps issued between host checks of the termination flag

// tiles resident at once: CUs x resident workgroups per CU x tiles per workgroup
long long tile_slots(const ts_ctx* c, int tiles_per_block) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->cfg.device) != hipSuccess || !c->ks.socc) return 0;
    return (long long)prop.multiProcessorCount * c->ks.socc() * tiles_per_block;
}

}  // namespace

int init_device_ctx(ts_ctx* c, const ts_config* cfg) {
    CHK(hipSetDevice(cfg->device));
    if (!c->arena) CHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const ts_config& k = c->cfg;
    GridArgs& g = c->grid;
    g.nd = c->nd;
    for (int a = 0; a < 3; a++) {
        g.n[a] = c->n[a];
        g.periodic[a] = a < c->nd ? k.periodic[a] : 1;
    }
    g.gamma = k.gamma;
    g.eps = k.eps;
    g.cfl = k.cfl;
    bool uniform = true;
    for (int a = 0; a < 3; a++) {
        std::vector<double> rc, rl, rr, ih;
        if (a < c->nd) {
            tsr::build_axis(*cfg, a, c->X[a], c->H[a]);
            uniform = uniform && tsr::uniform_axis(c->H[a]);
            int n = c->n[a];
            rc.resize(n);
            rl.resize(n);
            rr.resize(n);
            const std::vector<double>& X = c->X[a];
            for (int i = 0; i < n; i++) {
                rc[i] = 1.0 / (X[i + 2] - X[i]);
                rl[i] = 1.0 / (X[i + 1] - X[i]);
                rr[i] = 1.0 / (X[i + 2] - X[i + 1]);
            }
        } else {
            c->X[a].assign(1, 0.0);
            c->H[a].assign(1, 1.0);  // grid.h:343-344
            rc.assign(1, 0.0);
            rl.assign(1, 0.0);
            rr.assign(1, 0.0);
        }
        for (double hv : c->H[a]) ih.push_back(1.0 / hv);
        int r;
        if ((r = upload(c, c->H[a].data(), c->H[a].size(), &g.h[a]))) return r;
        if ((r = upload(c, ih.data(), ih.size(), &g.ih[a]))) return r;
        if ((r = upload(c, rc.data(), rc.size(), &g.rc[a]))) return r;
        if ((r = upload(c, rl.data(), rl.size(), &g.rl[a]))) return r;
        if ((r = upload(c, rr.data(), rr.size(), &g.rr[a]))) return r;
    }
    g.eps2 = 0.0;
    for (int a = 0; a < 3; a++) g.ev[a] = nullptr;
    if (k.recon == TS_RECON_LINEAR_LIM && k.limiter == TS_LIMITER_VENKAT) {
        // limVenkatar::SendCellInfo (Limiters.h:93-96) with K = 5 (Constants.h:11): one value on
        // a uniform grid, per-axis factors of the per-cell value otherwise
        g.eps2 = tsr::venkat_eps2(c->nd, c->H[0][1], c->nd > 1 ? c->H[1][1] : 1.0, c->nd > 2 ? c->H[2][1] : 1.0);
        if (!uniform) {
            for (int a = 0; a < 3; a++) {
                std::vector<double> f;
                tsr::venkat_axis_factors(c->nd, a < c->nd ? c->n[a] : 1, a < c->nd ? c->H[a].data() + 1 : nullptr, f);
                int r;
                if ((r = upload(c, f.data(), f.size(), &g.ev[a]))) return r;
            }
        }
    }
    for (int m = 0; m < TS_MAX_BC; m++) {
        const ts_bc_config& b = k.bc[m];
        g.bc[m].type = m < k.n_bc ? b.type : TS_BC_NATURAL;
        for (int a = 0; a < 3; a++) g.bc[m].vel[a] = b.velocity[a];
        g.bc[m].p = b.pstatic;
        g.bc[m].rho = b.density;
    }
    for (int s = 0; s < 6; s++) {
        g.side_marker[s] = k.side_marker[s];
        g.face_marker[s] = nullptr;
        int a = s / 2;
        if (a < c->nd && cfg->face_marker[s]) {
            size_t nf = c->ncell / c->n[a];
            for (size_t f = 0; f < nf; f++)
                if (cfg->face_marker[s][f] >= k.n_bc) {
                    c->err = "face marker refers to a missing boundary condition";
                    return TS_ERR_INVALID;
                }
            c->hmarker[s].assign(cfg->face_marker[s], cfg->face_marker[s] + nf);
            const uint8_t* d = nullptr;
            int r = upload(c, cfg->face_marker[s], nf, &d);
            if (r) return r;
            g.face_marker[s] = d;
        }
    }
    if (!tsr::lookup_kernels(c->nd, k.riemann, tsr::recon_variant(k), k.precision, &c->ks) || !c->ks.dt) {
        c->err = "no kernel instance for this (dimension, flux, reconstruction) combination";
        return TS_ERR_UNSUPPORTED;
    }
    c->vis = k.viscous || k.external_force || k.external_accel || k.viscosity != 0.0;
    if (c->vis) {
        if (!c->ks.vdt || !c->ks.vpre) {
            c->err = "no viscous kernel instance for this combination";
            return TS_ERR_UNSUPPORTED;
        }
        VisArgs& v = c->visargs;
        v.mu = k.viscosity;
        v.viscous = k.viscous;
        v.force = k.external_force;
        v.accel_on = k.external_accel;
        for (int a = 0; a < 3; a++) {
            v.sigma[a] = k.sigma[a];
            v.accel[a] = k.accel[a];
        }
        if (k.viscous) {
            int r;
            if ((r = dalloc(c, &v.grad, c->ncell * 3 * c->nd))) return r;
            for (int a = 0; a < c->nd; a++) {
                if ((r = upload(c, c->X[a].data(), c->X[a].size(), &v.x[a]))) return r;
                v.nface[a] = (unsigned)(c->ncell / c->n[a] * (c->n[a] + 1));
                if ((r = dalloc(c, &v.fv[a], (size_t)v.nface[a] * 4))) return r;
            }
        }
    }
    // state buffers
    int r;
    if (!c->dU[0]) {  // (borrowed-memory mode: the caller's buffer)
        double* u0 = nullptr;
        if ((r = dalloc(c, &u0, c->ncell * 5))) return r;
        c->dU[0] = u0;
    }
    c->stage_elem = (k.rk_stages == 3 && k.precision == TS_FP32_RESIDUAL) ? 4 : 8;
    // SSP-RK3: L(U^n) / U1 and U2; the FP32-residual form adds the stage-2 increment
    int nstage = k.rk_stages == 3 ? (c->stage_elem == 4 ? 3 : 2) : 1;
    for (int s = 1; s <= nstage; s++) {
        double* p = nullptr;
        size_t count = c->stage_elem == 8 ? c->ncell * 5 : (c->ncell * 5 + 1) / 2;
        if ((r = dalloc(c, &p, count))) return r;
        c->dU[s] = p;
    }
    if ((r = dalloc(c, &c->dctl, 1))) return r;
    if (c->arena) {
        c->hctl = &c->hctl_local;
        c->use_graphs = false;  // instantiating a graph allocates
    } else {
        CHK(hipHostMalloc((void**)&c->hctl, sizeof(StepCtl), hipHostMallocDefault));
    }
    std::memset(c->hctl, 0, sizeof(StepCtl));
    // tiling
    if (c->nd == 1) {
        c->ty = 1;
        c->nstrips = (c->n[0] + STRIP - 1) / STRIP;
        c->ntiles = c->nstrips;
        // four strips (one per wavefront) per block
        c->block_dim = dim3(256);
        c->grid_dim = dim3((c->ntiles + 3) / 4);
    } else if (c->nd == 2) {
        c->ty = tsr::tuned_rows_per_tile(*cfg, tile_slots(c, 4));
        c->nstrips = (c->n[0] + STRIP - 1) / STRIP;
        int nyb = (c->n[1] + c->ty - 1) / c->ty;
        c->ntiles = c->nstrips * nyb;
        c->block_dim = dim3(256);
        c->grid_dim = dim3((c->ntiles + 3) / 4);
    } else {
        c->ty = tsr::tuned_rows_per_tile(*cfg, tile_slots(c, 1));
        c->nstrips = (c->n[0] + STRIP - 1) / STRIP;
        int nyb = (c->n[1] + c->ty - 1) / c->ty;
        int nzb = (c->n[2] + ZOUT - 1) / ZOUT;
        c->ntiles = c->nstrips * nyb * nzb;
        c->block_dim = dim3(64 * ZW);
        c->grid_dim = dim3(c->ntiles);
    }
    // one minimum of V/SR per tile from every dt pass and fused stage 1
    if ((r = dalloc(c, &c->dt_part, (size_t)c->ntiles))) return r;
    const char* sep = getenv("TS_SEPARATE_DT");
    c->fused_dt = k.rk_stages == 3 && !(sep && sep[0] == '1') && c->ks.f1 && c->ks.f2 && !c->vis;
    c->euler1d_wg = k.rk_stages == 1 && c->nd == 1 && k.precision == TS_FP64 && !c->vis && c->ks.euler1d_chunk &&
                    c->ntiles <= 64 && !(sep && sep[0] == '1');
    return TS_OK;
}

void destroy_device_ctx(ts_ctx* c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; i++)
        if (c->graph_exec[i]) (void)hipGraphExecDestroy(c->graph_exec[i]);
    for (double* p : c->dbufs) (void)hipFree(p);
    c->dbufs.clear();
    if (c->dt_log) (void)hipFree(c->dt_log);
    if (c->rn_part) (void)hipFree(c->rn_part);
    if (c->d_cells) (void)hipFree(c->d_cells);
    if (c->d_cell_vals) (void)hipFree(c->d_cell_vals);
    if (c->d_aos) (void)hipFree(c->d_aos);
    if (c->hctl && c->hctl != &c->hctl_local) (void)hipHostFree(c->hctl);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    if (c->stream && !c->arena) (void)hipStreamDestroy(c->stream);
}

// The ghost layers of an AoS state whose interior is in place and whose ghost entries are zero:
// periodic images or boundary-condition images of the interior (kernel.h:334-915), one face
// layer per active axis (edge and corner ghosts stay zero).
void fill_ghost_images(const ts_ctx* c, double* aos) {
    int g[3] = {1, c->nd > 1 ? 1 : 0, c->nd > 2 ? 1 : 0};
    // ghost layers: periodic images or boundary-condition images (kernel.h:334-915)
    for (int a = 0; a < c->nd; a++) {
        int ja = a == 0 ? 1 : 0, ka = a == 2 ? 1 : 2;
        int nj = c->n[ja], nk = c->n[ka];
        for (int kk = 0; kk < nk; kk++)
            for (int jj = 0; jj < nj; jj++)
                for (int side = 0; side < 2; side++) {
                    int I[3], S[3];
                    I[ja] = S[ja] = jj + g[ja];
                    I[ka] = S[ka] = kk + g[ka];
                    int n = c->n[a];
                    if (c->cfg.periodic[a]) {
                        I[a] = side ? n + 1 : 0;
                        S[a] = side ? 1 : n;
                        for (int v = 0; v < 5; v++) aos[gidx(c, I[0], I[1], I[2]) * 5 + v] = aos[gidx(c, S[0], S[1], S[2]) * 5 + v];
                    } else {
                        I[a] = side ? n + 1 : 0;
                        S[a] = side ? n : 1;
                        int s = 2 * a + side;
                        int face = a == 0 ? kk * c->n[1] + jj : kk * c->n[0] + jj;
                        int m = c->hmarker[s].empty() ? c->cfg.side_marker[s] : c->hmarker[s][face];
                        tsr::host_bc_image(c->cfg.bc[m], c->cfg.gamma, a, &aos[gidx(c, S[0], S[1], S[2]) * 5],
                                           &aos[gidx(c, I[0], I[1], I[2]) * 5]);
                    }
                }
    }
}

namespace {

struct Launch {
    LaunchFn fn;
    KArgs args;
    int kind;
};

constexpr int MAX_LAUNCHES = 8;  // per step: 3 viscous pre-passes, dt, norms, 3 stages

// the launches of step s of a chunk, given the buffer holding U^n at the chunk start
void step_launches(const ts_ctx* c, int start_buf, int s, Launch (&L)[MAX_LAUNCHES], int& n) {
    KArgs a{};
    a.g = c->grid;
    a.v = c->visargs;
    a.plane = c->ncell;
    a.ctl = c->dctl;
    a.ty = c->ty;
    a.nstrips = c->nstrips;
    a.ntiles = c->ntiles;
    a.un = (const double*)c->dU[0];
    a.rn_part = c->rn_part;
    a.dt_part = c->dt_part;
    n = 0;
    // residual norms of U^n, after the kernel that decides whether the step runs
    auto norms = [&](const void* un_buf) {
        if (!c->norms) return;
        KArgs r = a;
        r.uin = un_buf;
        r.lin = nullptr;
        r.uout = nullptr;
        L[n++] = {c->vis ? c->ks.vrn : c->ks.rn, r, 5};
    };
    // viscous pre-pass over the stage input: while steps remain (before the dt pass), or when the
    // step runs (later stages)
    auto prepass = [&](const void* in, int gate_go) {
        if (!c->cfg.viscous) return;
        KArgs p = a;
        p.uin = in;
        p.v.gate_go = gate_go;
        L[n++] = {c->ks.vpre, p, 6};
    };
    if (c->vis) {
        if (c->cfg.rk_stages == 1) {
            int in = (start_buf + s) & 1, out = in ^ 1;
            a.uin = c->dU[in];
            prepass(a.uin, 0);
            L[n++] = {c->ks.vdt, a, 0};
            norms(c->dU[in]);
            a.uout = c->dU[out];
            L[n++] = {c->ks.vs1, a, 1};
        } else {
            a.uin = c->dU[0];
            prepass(a.uin, 0);
            L[n++] = {c->ks.vdt, a, 0};
            norms(c->dU[0]);
            a.uout = c->dU[1];
            L[n++] = {c->ks.vs1, a, 1};
            a.uin = c->dU[1];
            prepass(a.uin, 1);
            a.uout = c->dU[2];
            L[n++] = {c->ks.vs2, a, 2};
            a.uin = c->dU[2];
            prepass(a.uin, 1);
            a.uout = c->dU[0];
            L[n++] = {c->ks.vs3, a, 3};
        }
        return;
    }
    if (c->cfg.rk_stages == 1) {
        int in = (start_buf + s) & 1, out = in ^ 1;
        a.uin = c->dU[in];
        a.uout = nullptr;
        L[n++] = {c->ks.dt, a, 0};
        norms(c->dU[in]);
        a.uout = c->dU[out];
        L[n++] = {c->ks.s1e, a, 1};
    } else if (c->fused_dt) {
        a.uin = c->dU[0];
        a.uout = c->dU[1];
        L[n++] = {c->ks.f1, a, 1};  // L(U^n) and dt
        norms(c->dU[0]);
        a.lin = c->dU[1];
        a.uout = c->dU[2];
        a.dout = c->dU[3];  // FP32-residual form: the increment U2 - U^n (null in FP64)
        L[n++] = {c->ks.f2, a, 2};  // U2 from U^n + dt L(U^n) on the fly
        a.lin = nullptr;
        a.dout = nullptr;
        a.din = c->dU[3];
        a.uin = c->dU[2];
        a.uout = c->dU[0];
        L[n++] = {c->ks.s3, a, 3};
        a.din = nullptr;
    } else {
        a.uin = c->dU[0];
        L[n++] = {c->ks.dt, a, 0};
        norms(c->dU[0]);
        a.uout = c->dU[1];
        L[n++] = {c->ks.s1r, a, 1};
        a.uin = c->dU[1];
        a.uout = c->dU[2];
        L[n++] = {c->ks.s2, a, 2};
        a.uin = c->dU[2];
        a.uout = c->dU[0];
        L[n++] = {c->ks.s3, a, 3};
    }
}

int issue_direct(ts_ctx* c, int start_buf, int first, int count) {
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> rec;
    size_t need = c->timing ? (size_t)count * 2 * MAX_LAUNCHES : 0;
    while (c->events.size() < need) {
        hipEvent_t e;
        CHK(hipEventCreate(&e));
        c->events.push_back(e);
    }
    size_t ev = 0;
    for (int s = first; s < first + count; s++) {
        Launch L[MAX_LAUNCHES];
        int n;
        step_launches(c, start_buf, s, L, n);
        for (int i = 0; i < n; i++) {
            if (c->timing) {
                hipEvent_t e0 = c->events[ev++], e1 = c->events[ev++];
                CHK(hipEventRecord(e0, c->stream));
                L[i].fn(L[i].args, c->grid_dim, c->block_dim, c->stream);
                CHK(hipEventRecord(e1, c->stream));
                rec.push_back({L[i].kind, {e0, e1}});
            } else {
                L[i].fn(L[i].args, c->grid_dim, c->block_dim, c->stream);
            }
        }
    }
    CHK(hipGetLastError());
    if (c->timing) {
        CHK(hipStreamSynchronize(c->stream));
        for (auto& r : rec) {
            float ms = 0;
            CHK(hipEventElapsedTime(&ms, r.second.first, r.second.second));
            c->t_ms[r.first] += ms;
            c->t_n[r.first] += 1;
        }
    }
    return TS_OK;
}

int issue_graph(ts_ctx* c, int start_buf, int count) {
    int gi = start_buf & 1;
    if (!c->graph_exec[gi] || c->graph_steps[gi] != count) {
        if (c->graph_exec[gi]) {
            CHK(hipGraphExecDestroy(c->graph_exec[gi]));
            c->graph_exec[gi] = nullptr;
        }
        hipGraph_t graph;
        CHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < count; s++) {
            Launch L[MAX_LAUNCHES];
            int n;
            step_launches(c, start_buf, s, L, n);
            for (int i = 0; i < n; i++) L[i].fn(L[i].args, c->grid_dim, c->block_dim, c->stream);
        }
        CHK(hipStreamEndCapture(c->stream, &graph));
        hipError_t e = hipGraphInstantiate(&c->graph_exec[gi], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            c->err = tsr::hip_msg(e, "hipGraphInstantiate");
            c->graph_exec[gi] = nullptr;
            return TS_ERR_HIP;
        }
        c->graph_steps[gi] = count;
    }
    CHK(hipGraphLaunch(c->graph_exec[gi], c->stream));
    return TS_OK;
}

}  // namespace

namespace {
__global__ void __launch_bounds__(256) gather_cells(const double* __restrict__ u, size_t plane,
                                                    const unsigned* __restrict__ idx, int n,
                                                    double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 5) return;
    const int c = t / 5, v = t - c * 5;
    out[t] = u[(size_t)v * plane + idx[c]];
}
}  // namespace

int get_cells(ts_ctx* c, const int* ijk, size_t n, double* out) {
    if (n == 0) return TS_OK;
    std::vector<unsigned> idx(n);
    const int lo[3] = {1, c->nd > 1 ? 1 : 0, c->nd > 2 ? 1 : 0};
    for (size_t q = 0; q < n; q++) {
        int p[3];
        for (int a = 0; a < 3; a++) {
            p[a] = ijk[3 * q + a] - lo[a];
            if (p[a] < 0 || p[a] >= c->n[a]) {
                c->err = "cell index outside the interior";
                return TS_ERR_INVALID;
            }
        }
        idx[q] = (unsigned)(((size_t)p[2] * c->n[1] + p[1]) * c->n[0] + p[0]);
    }
    if (n > c->cells_cap) {
        if (c->d_cells) CHK(hipFree(c->d_cells));
        if (c->d_cell_vals) CHK(hipFree(c->d_cell_vals));
        c->d_cells = nullptr;
        c->d_cell_vals = nullptr;
        c->cells_cap = 0;
        CHK(hipMalloc((void**)&c->d_cells, n * sizeof(unsigned)));
        CHK(hipMalloc((void**)&c->d_cell_vals, n * 5 * sizeof(double)));
        c->cells_cap = n;
    }
    CHK(hipMemcpyAsync(c->d_cells, idx.data(), n * sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
    const int work = (int)(n * 5);
    hipLaunchKernelGGL(gather_cells, dim3((work + 255) / 256), dim3(256), 0, c->stream, (const double*)c->dU[c->cur],
                       (size_t)c->ncell, c->d_cells, (int)n, c->d_cell_vals);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(out, c->d_cell_vals, n * 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    return TS_OK;
}

int run_steps(ts_ctx* c, int max_steps, double t_max, double next_snap) {
    StepCtl* h = c->hctl;
    std::memset(h, 0, sizeof(StepCtl));
    h->time = c->h_time;
    h->dt = c->h_dt;
    h->t_max = t_max;
    h->next_snap = next_snap;
    h->iter = c->h_iter;
    h->max_steps = max_steps;
    h->steps_done = 0;
    h->active = max_steps > 0 ? 1 : 0;
    h->go = 0;
    h->error = 0;
    h->dt_log = c->dt_log;
    h->dt_log_base = c->dt_log_base;
    h->dt_log_cap = c->dt_log ? c->dt_log_cap : 0;
    CHK(hipMemcpyAsync(c->dctl, h, sizeof(StepCtl), hipMemcpyHostToDevice, c->stream));
    int launched = 0;
    const int start_buf = c->cur;
    int rc = TS_OK;
    const bool workgroup_chunks = c->euler1d_wg && !c->norms;
    if (workgroup_chunks) {
        // chunks of steps in one workgroup; the state buffer of the next chunk follows the steps
        // taken so far
        constexpr int WG_CHUNK = 4096;
        KArgs a{};
        a.g = c->grid;
        a.plane = c->ncell;
        a.ctl = c->dctl;
        a.ty = c->ty;
        a.nstrips = c->nstrips;
        a.ntiles = c->ntiles;
        int done = 0;
        while (launched < max_steps) {
            const int chunk = std::min(max_steps - launched, WG_CHUNK);
            const int in = (start_buf + done) & 1;
            a.uin = c->dU[in];
            a.uout = c->dU[in ^ 1];
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (c->timing) {
                while (c->events.size() < 2) {
                    hipEvent_t e;
                    CHK(hipEventCreate(&e));
                    c->events.push_back(e);
                }
                e0 = c->events[0];
                e1 = c->events[1];
                CHK(hipEventRecord(e0, c->stream));
            }
            c->ks.euler1d_chunk(a, chunk, c->stream);
            CHK(hipGetLastError());
            if (c->timing) CHK(hipEventRecord(e1, c->stream));
            launched += chunk;
            CHK(hipMemcpyAsync(h, c->dctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
            CHK(hipStreamSynchronize(c->stream));
            if (c->timing) {
                float ms = 0;
                CHK(hipEventElapsedTime(&ms, e0, e1));
                c->t_ms[1] += ms;
                c->t_n[1] += 1;
            }
            done = h->steps_done;
            if (!h->active) break;
        }
    }
    while (!workgroup_chunks && launched < max_steps) {
        int chunk = std::min(max_steps - launched, CHUNK);
        // executed steps always form a prefix of the issued ones, so the Euler ping-pong buffer
        // of the final state is (start + steps_done) mod 2 whatever the chunking
        int sb = (start_buf + launched) & 1;
        if (c->use_graphs && !c->timing && chunk == CHUNK) rc = issue_graph(c, sb, chunk);
        else rc = issue_direct(c, sb, 0, chunk);
        if (rc) return rc;
        launched += chunk;
        CHK(hipMemcpyAsync(h, c->dctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
        CHK(hipStreamSynchronize(c->stream));
        if (!h->active) break;
    }
    if (launched == 0) {
        CHK(hipMemcpyAsync(h, c->dctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
        CHK(hipStreamSynchronize(c->stream));
    }
    c->h_time = h->time;
    c->h_dt = h->dt;
    c->h_iter = h->iter;
    c->h_steps_done = h->steps_done;
    for (int v = 0; v < 5; v++) c->h_rnorm[v] = c->norms ? h->rnorm[v] : 0.0;
    if (c->cfg.rk_stages == 1) c->cur = (start_buf + h->steps_done) & 1;
    if (h->error) {
        c->err = "dt is not finite and positive (non-physical state)";
        return TS_ERR_UNPHYSICAL;
    }
    return TS_OK;
}

