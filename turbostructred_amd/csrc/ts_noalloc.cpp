// ts_noalloc.cpp -- ts_step_noalloc: the step loop of ts_step for callers that cannot allocate
// device memory.  Its device buffers (geometry tables, control block, stage arrays) come from one
// fixed region that the library owns for the life of the process on each GPU it runs on.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ts_internal.hpp"

namespace {

// The fixed region.  Its name is a host-side stand-in for the device symbol: host code uses
// only the per-device address obtained from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[TS_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // per-device address of `scratch`, set on first use

// device address of the region on the current device `dev` (looked up once per device)
int scratch_address(int dev, void** out, std::string& err) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = tsr::hip_msg(e, "hipGetSymbolAddress(scratch)");
            return TS_ERR_HIP;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return TS_OK;
}

}  // namespace

extern "C" int ts_step_noalloc(const ts_config* cfg, double* d_state, double time, long long iteration,
                               int max_steps, double t_max, double next_snapshot_time, ts_step_info* out,
                               void* stream) {
    std::string err;
    int rc = tsr::validate(cfg, err);
    if (rc) {
        tsr::set_thread_error("ts_step_noalloc: " + err);
        return rc;
    }
    if (!d_state || max_steps < 0) {
        tsr::set_thread_error("ts_step_noalloc: null state or negative max_steps");
        return TS_ERR_INVALID;
    }
    if (cfg->device < 0 || cfg->device >= kMaxDevices) {
        tsr::set_thread_error("ts_step_noalloc: device index out of range");
        return TS_ERR_INVALID;
    }
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) {
        tsr::set_thread_error("ts_step_noalloc: " + tsr::hip_msg(e, "hipSetDevice"));
        return TS_ERR_HIP;
    }
    void* base = nullptr;
    if ((rc = scratch_address(cfg->device, &base, err))) {
        tsr::set_thread_error("ts_step_noalloc: " + err);
        return rc;
    }
    // a context on the stack whose buffers are carved out of the region; dU[0] is the caller's
    ts_ctx c;
    c.cfg = *cfg;
    for (int s = 0; s < 6; s++) c.cfg.face_marker[s] = nullptr;
    for (int a = 0; a < 3; a++) { c.cfg.coord[a] = nullptr; c.cfg.h[a] = nullptr; }
    c.nd = cfg->ndims;
    c.n[0] = cfg->nx;
    c.n[1] = cfg->ny;
    c.n[2] = cfg->nz;
    c.ncell = (size_t)c.n[0] * c.n[1] * c.n[2];
    c.arena = (char*)base;
    c.arena_cap = TS_NOALLOC_SCRATCH_BYTES;
    c.stream = (hipStream_t)stream;
    c.dU[0] = d_state;
    rc = init_device_ctx(&c, cfg);
    if (rc == TS_OK) {
        c.h_time = time;
        c.h_iter = iteration;
        rc = run_steps(&c, max_steps, t_max, next_snapshot_time);
        // forward Euler ends in the other buffer after an odd number of steps: copy it back
        if (c.cur != 0 && (rc == TS_OK || rc == TS_ERR_UNPHYSICAL)) {
            e = hipMemcpyAsync(d_state, c.dU[c.cur], c.ncell * 5 * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
            if (e != hipSuccess) {
                c.err = tsr::hip_msg(e, "hipMemcpyAsync");
                rc = TS_ERR_HIP;
            }
        }
        if (out) {
            out->time = c.h_time;
            out->dt = c.h_dt;
            out->iteration = c.h_iter;
            for (int v = 0; v < 5; v++) out->residual_norm[v] = 0.0;
            out->steps_done = c.h_steps_done;
        }
    }
    destroy_device_ctx(&c);  // frees nothing in this mode: synchronises the caller's stream
    if (rc) tsr::set_thread_error("ts_step_noalloc: " + c.err);
    else tsr::set_thread_error("");
    return rc;
}
