// ts_batch.cpp -- batching and serving: many independent solves advanced together on one GPU.
//
// Every resident solve keeps its own device state (U^n in double, L(U^n) and U^2 in the
// stage-array type), geometry and control block inside one device arena.  A step of the batch is
// the fused-dt SSP-RK3 sequence (three launches per kernel family); each launch is a persistent
// grid over a (problem, tile) table that spans all resident solves of that family; each solve takes its own dt, stops at its own t_end
// or step count, and its arithmetic does not depend on what else is in the batch.  Between steps
// the engine retires finished solves (their state is copied back) and admits queued ones that
// fit the HBM budget: continuous batching.  A background thread can drive it (serving).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "ts_internal.hpp"

using namespace tsd;

namespace {

constexpr size_t ALIGN = 256;

static_assert(sizeof(StepCtl) % 8 == 0, "control block is copied as 8-byte words");

// One launch packs the control blocks of every resident solve (scattered over the arena) into a
// contiguous buffer, so that they come back to the host in a single copy.
// descriptors uploaded contiguously, then placed at their slots: one block per descriptor
static_assert(sizeof(KArgs) % 8 == 0, "descriptors are copied in 8-byte words");
__global__ void __launch_bounds__(64) scatter_desc(const KArgs* __restrict__ src, const int* __restrict__ dst_index,
                                                   int n, KArgs* __restrict__ dst) {
    const int i = blockIdx.x;
    if (i >= n) return;
    constexpr int W = sizeof(KArgs) / 8;
    const unsigned long long* s = reinterpret_cast<const unsigned long long*>(src + i);
    unsigned long long* d = reinterpret_cast<unsigned long long*>(dst + dst_index[i]);
    for (int w = threadIdx.x; w < W; w += blockDim.x) d[w] = s[w];
}

__global__ void __launch_bounds__(256) gather_ctl(StepCtl* const* src, StepCtl* dst, int n) {
    constexpr int W = sizeof(StepCtl) / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * W) return;
    const int p = i / W, w = i - p * W;
    reinterpret_cast<unsigned long long*>(dst + p)[w] = reinterpret_cast<const unsigned long long*>(src[p])[w];
}
inline size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// First-fit free list over one device allocation.
class Arena {
   public:
    size_t capacity = 0;
    char* base = nullptr;
    std::map<size_t, size_t> free_;  // offset -> size
    size_t used = 0;
    void init(char* b, size_t cap) {
        base = b;
        capacity = cap;
        free_.clear();
        free_[0] = cap;
        used = 0;
    }
    bool alloc(size_t n, size_t* off) {
        n = align_up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                size_t rest = it->second - n;
                size_t o = it->first;
                free_.erase(it);
                if (rest) free_[o + n] = rest;
                used += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align_up(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

using FamilyKey = std::tuple<int, int, int, int>;  // ndims, riemann, recon variant, precision

struct Problem {
    long long ticket = 0;
    ts_config cfg{};
    int max_steps = 0;
    double t_end = 0;
    std::vector<double> init;  // reference AoS state until admission
    std::vector<double> coord[3], h[3];  // owned copies of optional non-uniform grid arrays
    size_t bytes = 0, off = 0;
    KArgs desc{};
    int family = -1;
    int ntiles = 0;
    size_t ncell = 0;
    int n[3] = {1, 1, 1};
    std::vector<double> result;
    ts_step_info info{};
    int status = 0;  // 0 queued, 1 resident (or being admitted), 2 done
    int error = 0;
    int steps_seen = 0;  // steps done as of the last pass (the autotuner's work count)
    int slot = -1;       // index of its descriptors in its family's descriptor blocks
    StepCtl final_ctl{};  // its control block when it finished (published with its state)
    bool uploaded = false;  // its descriptors are in place for the family's current tiling
    // admission: the host sources of its asynchronous uploads, kept until the copy stream is
    // synchronised
    std::vector<double> geo_host;
    StepCtl init_ctl{};
};

// Tile-shape autotuning of a family (ts_batch_config.autotune_tiles): for each load bucket
// (floor(log2) of the family's resident cells) every candidate tiling is timed over whole
// passes, and the fastest per cell-step is kept.  The candidates take their passes in the
// palindromic order 0 1 2 3 3 2 1 0, so that a load that drifts steadily within the bucket (solves
// finishing) weighs on every candidate alike.  Results do not depend on the tiling.
constexpr int kTileCands = 4;
constexpr int kTrialPasses = 2;
constexpr int kTrialOrder[kTileCands * kTrialPasses] = {0, 1, 2, 3, 3, 2, 1, 0};
constexpr int kCand2D[kTileCands] = {4, 8, 16, 32};  // output rows per tile
constexpr int kCand1D[kTileCands] = {1, 2, 4, 8};    // 60-cell strips per work item

struct TileTrial {
    double ms = 0, cell_steps = 0;
    int passes = 0;
};
struct LoadTuning {
    TileTrial t[kTileCands];
    int chosen = -1;
};

struct Family {
    FamilyKey key;
    KernelSet ks;
    int occupancy = 1;
    // device tables
    KArgs* d_desc = nullptr;
    int2* d_table = nullptr;
    int* d_ntiles = nullptr;  // tiles in the table (read by the kernels)
    size_t cap_desc = 0, cap_table = 0;
    // stage 1's minimum of V/SR per table entry, and per member its slot, first entry and entry
    // count (the members' entries are contiguous in the table), read by dt_finalize_batched
    double* d_part = nullptr;
    int4* d_ranges = nullptr;
    int* d_nmem = nullptr;
    size_t cap_ranges = 0;
    int h_nmem = 0;
    // resident problems in descriptor order.  A Problem is owned by ts_batch::probs through a
    // unique_ptr, so its address is stable: the run thread reads members without the map (and
    // without the lock), while submit inserts into and poll erases from the map under the lock.
    std::vector<Problem*> members;
    bool dirty = true;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    // descriptor slots: a resident problem keeps its slot until it retires, so a membership change
    // uploads only the new problems' descriptors (all of them when the tiling or the arrays change)
    std::vector<int> free_slots;
    int slot_end = 0;        // slots handed out so far: [0, slot_end)
    int param = -1;          // tiling parameter chosen by the last retile (rows, or strips per item)
    int uploaded_param = -1; // the one the uploaded descriptors carry
    KArgs* d_fresh = nullptr;  // staging for new descriptors and their destination indices
    int* d_fresh_idx = nullptr;
    size_t cap_fresh = 0;
    int h_ntiles = 0;
    // autotuning: trials per load bucket, the tiling in use, and the timing of the current pass
    std::map<int, LoadTuning> tuning;
    int bucket = -1, cand = -1;  // cand -1: the load rule
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;
    double pass_cell_steps = 0;  // cell-steps the timed pass actually advanced
};

}  // namespace

// Locking.  `mu` guards the queue, the ticket map, the arena and the family list's length; it
// is held only for bookkeeping, never across a device copy or synchronisation.  `run_mu` is held
// for the whole of one admission + advance pass (batch_run), by the server thread and by
// ts_batch_run alike, so that two passes never share the stream, the control-block buffers or a
// family's tables and graph.  The run thread reads Family::members and the Problems in it
// without `mu`: only a pass (under run_mu) adds or removes members, and poll only erases
// problems that have left every family.
struct ts_batch {
    ts_batch_config bc{};
    std::string err;
    std::mutex err_mu;  // err is written by the caller's thread and by the server thread
    std::mutex mu;
    std::mutex run_mu;
    std::mutex server_mu;  // serialises ts_batch_start_server / ts_batch_stop_server
    std::condition_variable cv;
    hipStream_t stream = nullptr;
    Arena arena;
    char* d_arena = nullptr;
    int ncu = 256;
    long long next_ticket = 1;
    std::deque<long long> queue;
    std::unordered_map<long long, std::unique_ptr<Problem>> probs;
    // capacity kMaxFamilies is reserved at creation and never exceeded: a submit that adds a
    // family (under the lock) then never moves the ones the server thread works on without it
    std::vector<Family> families;
    double* d_stage = nullptr;  // device staging of AoS states for the layout conversion (uploads)
    size_t stage_cap = 0;       // doubles
    // Finished solves leave through a second stream: their states are converted and copied out
    // while the next pass's kernels run, and published (done, memory released) at that pass.
    hipStream_t cstream = nullptr;
    double* d_rstage = nullptr;  // staging of the outgoing AoS states, one region per solve
    size_t rstage_cap = 0;
    std::vector<Problem*> retiring;
    StepCtl* h_ctl = nullptr;   // pinned, one per resident problem
    StepCtl** h_ptr = nullptr;  // pinned, device address of each resident problem's control block
    StepCtl* d_ctl = nullptr;   // packed copies of the resident control blocks
    StepCtl** d_ptr = nullptr;
    size_t h_ctl_cap = 0;
    std::thread server;
    std::atomic<bool> serving{false};
    std::atomic<bool> stop{false};
    // steps per server pass: the C5 stream at 200 requests/s had p50 153 / 125 / 104 ms with 4 / 8
    // / 16 steps (fewer host round trips per step) and the same throughput at 2000 requests/s
    int steps_per_chunk = 16;
};

namespace {

int fail(ts_batch* b, int code, const std::string& m) {
    std::lock_guard<std::mutex> lk(b->err_mu);
    b->err = m;
    return code;
}

#define BCHK(call)                                      \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) {                         \
            return fail(b, TS_ERR_HIP, tsr::hip_msg(e_, #call)); \
        }                                               \
    } while (0)

size_t state_cells(const ts_config& c) { return (size_t)c.nx * c.ny * c.nz; }

// geometry + state + control of one problem laid out in its arena block
struct Layout {
    size_t u0, l, u2, d2, geo, ctl, markers, total;  // d2: FP32-residual stage-2 increment
    size_t geo_doubles;
};

Layout layout_of(const ts_config& c) {
    Layout L{};
    size_t n = state_cells(c);
    size_t se = c.precision == TS_FP32_RESIDUAL ? 4 : 8;
    size_t geo = 0;
    int na[3] = {c.nx, c.ny, c.nz};
    for (int a = 0; a < 3; a++) {
        int m = a < c.ndims ? na[a] : 1;
        // h, ih (with ghosts); rc, rl, rr and the Venkatakrishnan factors (interior)
        geo += 2 * (m + 2) + 4 * m;
    }
    L.geo_doubles = geo;
    size_t o = 0;
    L.u0 = o; o = align_up(o + n * 5 * 8);
    L.l = o; o = align_up(o + n * 5 * se);
    L.u2 = o; o = align_up(o + n * 5 * se);
    L.d2 = o; o = align_up(o + (se == 4 ? n * 5 * se : 0));
    L.geo = o; o = align_up(o + geo * 8);
    L.ctl = o; o = align_up(o + sizeof(StepCtl));
    L.markers = o;
    size_t mk = 0;
    for (int s = 0; s < 6; s++)
        if (s / 2 < c.ndims && c.face_marker[s]) mk += n / na[s / 2];
    o = align_up(o + mk);
    L.total = o;
    return L;
}

// every (ndims, flux, reconstruction variant, precision) key: 3 x 3 x 5 x 2
constexpr size_t kMaxFamilies = 90;

int family_of(ts_batch* b, const ts_config& c) {
    FamilyKey k(c.ndims, c.riemann, tsr::recon_variant(c), c.precision);
    for (size_t i = 0; i < b->families.size(); i++)
        if (b->families[i].key == k) return (int)i;
    Family f;
    f.key = k;
    if (!tsr::lookup_kernels(c.ndims, c.riemann, tsr::recon_variant(c), c.precision, &f.ks) || !f.ks.bf1) return -1;
    f.occupancy = f.ks.bocc ? f.ks.bocc() : 1;
    if (b->families.size() >= kMaxFamilies) return -1;
    b->families.push_back(f);
    return (int)b->families.size() - 1;
}

// the device staging buffer holds at least `doubles` values (only the pass thread uses it)
int ensure_stage(ts_batch* b, size_t doubles) {
    if (doubles <= b->stage_cap) return TS_OK;
    BCHK(hipStreamSynchronize(b->cstream));  // earlier uploads may still read the old buffer
    if (b->d_stage) BCHK(hipFree(b->d_stage));
    b->d_stage = nullptr;
    b->stage_cap = 0;
    BCHK(hipMalloc((void**)&b->d_stage, doubles * sizeof(double)));
    b->stage_cap = doubles;
    return TS_OK;
}

// upload geometry/state/control of a problem into its arena block and fill its descriptor.  The
// copies and the layout conversion go on the copy stream, so that they overlap the kernels of
// a pass in flight; the caller synchronises that stream before the problem runs and before the
// host sources (init, geo_host, init_ctl) are released.
int admit(ts_batch* b, Problem& p) {
    const ts_config& c = p.cfg;
    Layout L = layout_of(c);
    char* base = b->d_arena + p.off;
    GridArgs& g = p.desc.g;
    g = GridArgs{};
    g.nd = c.ndims;
    g.n[0] = c.nx; g.n[1] = c.ny; g.n[2] = c.nz;
    for (int a = 0; a < 3; a++) g.periodic[a] = a < c.ndims ? c.periodic[a] : 1;
    g.gamma = c.gamma;
    g.eps = c.eps;
    g.cfl = c.cfl;
    std::vector<double>& geo = p.geo_host;
    geo.clear();
    geo.reserve(L.geo_doubles);
    double* dgeo = (double*)(base + L.geo);
    bool uniform = true;
    std::vector<double> X, H, EV;
    double h1[3] = {1.0, 1.0, 1.0};
    size_t oev[3] = {0, 0, 0};
    const bool venkat = c.recon == TS_RECON_LINEAR_LIM && c.limiter == TS_LIMITER_VENKAT;
    for (int a = 0; a < 3; a++) {
        int n = a < c.ndims ? p.n[a] : 1;
        if (a < c.ndims) {
            tsr::build_axis(c, a, X, H);
            uniform = uniform && tsr::uniform_axis(H);
        } else {
            X.assign(3, 0.0);
            H.assign(3, 1.0);
        }
        size_t oh = geo.size();
        h1[a] = H[1];
        for (int i = 0; i < n + 2; i++) geo.push_back(H[i]);
        size_t oi = geo.size();
        for (int i = 0; i < n + 2; i++) geo.push_back(1.0 / H[i]);
        size_t orc = geo.size();
        for (int i = 0; i < n; i++) geo.push_back(a < c.ndims ? 1.0 / (X[i + 2] - X[i]) : 0.0);
        size_t orl = geo.size();
        for (int i = 0; i < n; i++) geo.push_back(a < c.ndims ? 1.0 / (X[i + 1] - X[i]) : 0.0);
        size_t orr = geo.size();
        for (int i = 0; i < n; i++) geo.push_back(a < c.ndims ? 1.0 / (X[i + 2] - X[i + 1]) : 0.0);
        oev[a] = geo.size();
        tsr::venkat_axis_factors(c.ndims, n, a < c.ndims ? H.data() + 1 : nullptr, EV);
        for (int i = 0; i < n; i++) geo.push_back(EV.size() == 1 ? EV[0] : EV[i]);
        g.h[a] = dgeo + oh;
        g.ih[a] = dgeo + oi;
        g.rc[a] = dgeo + orc;
        g.rl[a] = dgeo + orl;
        g.rr[a] = dgeo + orr;
    }
    // limVenkatar::SendCellInfo (Limiters.h:93-96), K = 5: one value on a uniform grid, per-axis
    // factors of the per-cell value otherwise
    g.eps2 = venkat ? tsr::venkat_eps2(c.ndims, h1[0], h1[1], h1[2]) : 0.0;
    for (int a = 0; a < 3; a++) g.ev[a] = (venkat && !uniform) ? dgeo + oev[a] : nullptr;
    for (int m = 0; m < TS_MAX_BC; m++) {
        g.bc[m].type = m < c.n_bc ? c.bc[m].type : TS_BC_NATURAL;
        for (int a = 0; a < 3; a++) g.bc[m].vel[a] = c.bc[m].velocity[a];
        g.bc[m].p = c.bc[m].pstatic;
        g.bc[m].rho = c.bc[m].density;
    }
    for (int s = 0; s < 6; s++) g.side_marker[s] = c.side_marker[s];
    BCHK(hipMemcpyAsync(dgeo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice, b->cstream));
    // state: the AoS array with ghosts crosses as it is and is converted on the device (the
    // staging buffer is reused by the next admission in the stream's order)
    size_t n = p.ncell;
    int r = ensure_stage(b, p.init.size());
    if (r) return r;
    BCHK(hipMemcpyAsync(b->d_stage, p.init.data(), p.init.size() * 8, hipMemcpyHostToDevice, b->cstream));
    BCHK(tsr::device_aos_to_soa(b->d_stage, (double*)(base + L.u0), c.ndims, p.n, b->cstream));
    StepCtl& ctl = p.init_ctl;
    ctl = StepCtl{};
    ctl.t_max = p.t_end > 0 ? p.t_end : 1e300;
    ctl.next_snap = p.t_end > 0 ? p.t_end : 0.0;
    ctl.max_steps = p.max_steps;
    ctl.active = p.max_steps > 0 ? 1 : 0;
    BCHK(hipMemcpyAsync(base + L.ctl, &ctl, sizeof(StepCtl), hipMemcpyHostToDevice, b->cstream));
    KArgs& d = p.desc;
    d.uin = nullptr;
    d.lin = nullptr;
    d.un = (const double*)(base + L.u0);
    d.uout = nullptr;
    d.plane = n;
    d.ctl = (StepCtl*)(base + L.ctl);
    // tile height: set for the whole family when its tables are built (retile)
    d.ty = c.ndims == 1 ? 1 : 4;
    d.nstrips = (c.nx + STRIP - 1) / STRIP;
    d.ntiles = d.nstrips * (c.ndims == 1 ? 1 : (c.ny + d.ty - 1) / d.ty);
    p.ntiles = d.ntiles;
    p.status = 1;
    return TS_OK;
}

// descriptor of problem p for one kernel of the step
KArgs stage_desc(const Problem& p, int kernel, char* arena) {
    Layout L = layout_of(p.cfg);
    char* base = arena + p.off;
    KArgs a = p.desc;
    void* u0 = base + L.u0;
    void* l = base + L.l;
    void* u2 = base + L.u2;
    // FP32-residual: stage 2 writes the increment U2 - U^n beside U2, stage 3 accumulates in
    // double from it
    void* d2 = p.cfg.precision == TS_FP32_RESIDUAL ? (void*)(base + L.d2) : nullptr;
    if (kernel == 0) { a.uin = u0; a.uout = l; }
    else if (kernel == 1) { a.uin = u0; a.lin = l; a.uout = u2; a.dout = d2; }
    else { a.uin = u2; a.uout = u0; a.din = d2; }
    return a;
}

// Tiling of a family, chosen again whenever its membership changes (results do not depend on
// the tiling): the load rule in retile, or the autotuner's candidates below.
int next_candidate(const LoadTuning& lt) {
    if (lt.chosen >= 0) return lt.chosen;
    int done = 0;
    for (const TileTrial& t : lt.t) done += t.passes;
    return kTrialOrder[done < kTileCands * kTrialPasses ? done : kTileCands * kTrialPasses - 1];
}

// the candidate the autotuner runs next for the family's current load (-1: autotuning off)
int tuning_candidate(ts_batch* b, Family& f) {
    long long cells = 0;
    for (const Problem* q : f.members) cells += (long long)q->ncell;
    int bucket = 0;
    while (bucket < 62 && (1ll << (bucket + 1)) <= cells) bucket++;
    f.bucket = bucket;
    if (!b->bc.autotune_tiles) return -1;
    return next_candidate(f.tuning[bucket]);
}

void retile(ts_batch* b, Family& f) {
    const long long waves = (long long)b->ncu * f.occupancy * 4 / 2;
    f.cand = tuning_candidate(b, f);
    if (std::get<0>(f.key) == 1) {
        // 1-D: a work item is a group of G consecutive 60-cell strips.  One strip per item
        // unless the strips outnumber the wavefronts many times over (then G = 4 or 2, the
        // largest that still leaves 8 items per wavefront): an item costs a descriptor read and
        // a stored minimum for the dt, cheap next to the parallelism of small items (the
        // autotuner picks 1 for every load of the 4096-tube workload but the full one)
        int G = 1;
        for (int cand = 4; cand > 1; cand /= 2) {
            long long items = 0;
            for (const Problem* q : f.members) items += (q->desc.nstrips + cand - 1) / cand;
            if (items >= 16 * waves) {
                G = cand;
                break;
            }
        }
        if (f.cand >= 0) G = kCand1D[f.cand];
        f.param = G;
        for (Problem* q : f.members) {
            Problem& p = *q;
            p.desc.ty = G;
            p.desc.ntiles = p.desc.nstrips;
            p.ntiles = (p.desc.nstrips + G - 1) / G;
        }
        return;
    }
    // 2-D: the tile height that minimises rounds x (rows + 2 warm-up rows), as for a single
    // solve: the persistent wavefronts take the family's tiles in rounds of all of them, and a
    // tile's time goes as the rows it marches.  (Taking the tallest height that still gave half
    // the wavefronts a tile ran a lone 2048^2 solve at 32 rows, 2240 tiles for 3072 wavefronts:
    // 1.76x the single-solve step.)
    const long long slots = 2 * waves;
    int hi = 4;
    for (const Problem* q : f.members) hi = std::max(hi, std::min(q->n[1], 64));
    int ty = 4;
    double best = 1e300;
    for (int cand = 4; cand <= hi; cand++) {
        long long tiles = 0;
        for (const Problem* q : f.members) {
            const Problem& p = *q;
            tiles += (long long)p.desc.nstrips * ((p.n[1] + cand - 1) / cand);
        }
        const double cost = double((tiles + slots - 1) / slots) * (cand + 2);
        if (cost < best) {
            best = cost;
            ty = cand;
        }
    }
    if (f.cand >= 0) ty = kCand2D[f.cand];
    f.param = ty;
    for (Problem* q : f.members) {
        Problem& p = *q;
        p.desc.ty = ty;
        p.desc.ntiles = p.desc.nstrips * ((p.n[1] + ty - 1) / ty);
        p.ntiles = p.desc.ntiles;
    }
}

int rebuild_tables(ts_batch* b, Family& f) {
    retile(b, f);
    size_t nt = 0;
    for (const Problem* q : f.members) nt += q->ntiles;
    bool realloc = false;
    if ((size_t)f.slot_end > f.cap_desc) {  // cap_desc descriptors per kernel, three kernels
        if (f.d_desc) BCHK(hipFree(f.d_desc));
        f.d_desc = nullptr;
        f.cap_desc = std::max<size_t>((size_t)f.slot_end * 2, 64);
        BCHK(hipMalloc(&f.d_desc, 3 * f.cap_desc * sizeof(KArgs)));
        realloc = true;
    }
    if (nt > f.cap_table) {
        if (f.d_table) BCHK(hipFree(f.d_table));
        if (f.d_part) BCHK(hipFree(f.d_part));
        f.d_table = nullptr;
        f.d_part = nullptr;
        f.cap_table = std::max<size_t>(nt * 2, 4096);
        BCHK(hipMalloc(&f.d_table, f.cap_table * sizeof(int2)));
        BCHK(hipMalloc(&f.d_part, f.cap_table * sizeof(double)));
        realloc = true;
    }
    if (f.members.size() > f.cap_ranges) {
        if (f.d_ranges) BCHK(hipFree(f.d_ranges));
        f.d_ranges = nullptr;
        f.cap_ranges = std::max<size_t>(f.members.size() * 2, 256);
        BCHK(hipMalloc(&f.d_ranges, f.cap_ranges * sizeof(int4)));
        realloc = true;
    }
    if (!f.d_ntiles) {
        BCHK(hipMalloc(&f.d_ntiles, sizeof(int)));
        BCHK(hipMalloc(&f.d_nmem, sizeof(int)));
        realloc = true;
    }
    if (realloc && f.graph) {
        (void)hipGraphExecDestroy(f.graph);
        f.graph = nullptr;
    }
    // descriptors: kernel k's block starts at k * cap_desc, so that the launch arguments stay
    // fixed while the resident set changes; only problems whose descriptors are not in place for
    // the current tiling are uploaded, contiguously, and scattered to their slots on the device
    const bool all = realloc || f.param != f.uploaded_param;
    std::vector<KArgs> fresh;
    std::vector<int> fresh_idx;
    for (Problem* q : f.members) {
        if (!all && q->uploaded) continue;
        for (int k = 0; k < 3; k++) {
            fresh.push_back(stage_desc(*q, k, b->d_arena));
            fresh_idx.push_back(int(k * f.cap_desc) + q->slot);
        }
        q->uploaded = true;
    }
    f.uploaded_param = f.param;
    if (!fresh.empty()) {
        if (fresh.size() > f.cap_fresh) {
            if (f.d_fresh) BCHK(hipFree(f.d_fresh));
            if (f.d_fresh_idx) BCHK(hipFree(f.d_fresh_idx));
            f.d_fresh = nullptr;
            f.d_fresh_idx = nullptr;
            f.cap_fresh = std::max<size_t>(fresh.size() * 2, 192);
            BCHK(hipMalloc(&f.d_fresh, f.cap_fresh * sizeof(KArgs)));
            BCHK(hipMalloc(&f.d_fresh_idx, f.cap_fresh * sizeof(int)));
        }
        BCHK(hipMemcpyAsync(f.d_fresh, fresh.data(), fresh.size() * sizeof(KArgs), hipMemcpyHostToDevice, b->stream));
        BCHK(hipMemcpyAsync(f.d_fresh_idx, fresh_idx.data(), fresh_idx.size() * sizeof(int), hipMemcpyHostToDevice,
                            b->stream));
        hipLaunchKernelGGL(scatter_desc, dim3((unsigned)fresh.size()), dim3(64), 0, b->stream, f.d_fresh,
                           f.d_fresh_idx, (int)fresh.size(), f.d_desc);
        BCHK(hipGetLastError());
    }
    std::vector<int2> table;
    std::vector<int4> ranges;
    table.reserve(nt);
    ranges.reserve(f.members.size());
    for (const Problem* q : f.members) {
        ranges.push_back(make_int4(q->slot, (int)table.size(), q->ntiles, 0));
        for (int t = 0; t < q->ntiles; t++) table.push_back(make_int2(q->slot, t));
    }
    if (nt) BCHK(hipMemcpyAsync(f.d_table, table.data(), nt * sizeof(int2), hipMemcpyHostToDevice, b->stream));
    if (!ranges.empty())
        BCHK(hipMemcpyAsync(f.d_ranges, ranges.data(), ranges.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    f.h_ntiles = (int)nt;
    f.h_nmem = (int)ranges.size();
    BCHK(hipMemcpyAsync(f.d_ntiles, &f.h_ntiles, sizeof(int), hipMemcpyHostToDevice, b->stream));
    BCHK(hipMemcpyAsync(f.d_nmem, &f.h_nmem, sizeof(int), hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));  // the copy sources are local vectors
    f.dirty = false;
    return TS_OK;
}

}  // namespace

namespace {

// launches of one batched step for family f
void issue_step(ts_batch* b, Family& f) {
    int nblk = b->ncu * f.occupancy;
    f.ks.bf1(f.d_desc, f.d_table, f.d_ntiles, f.d_part, nblk, b->stream);
    // each member's dt from its entries' minima: one wavefront per member, 8 per CU
    hipLaunchKernelGGL(tsd::dt_finalize_batched<0>, dim3(2 * b->ncu), dim3(256), 0, b->stream, f.d_desc, f.d_ranges,
                       f.d_nmem, f.d_part);
    f.ks.bf2(f.d_desc + f.cap_desc, f.d_table, f.d_ntiles, f.d_part, nblk, b->stream);
    f.ks.bs3(f.d_desc + 2 * f.cap_desc, f.d_table, f.d_ntiles, f.d_part, nblk, b->stream);
}

// `timed`: events bracket the launches themselves (a graph is built before the first event)
int issue_steps(ts_batch* b, Family& f, int n, bool timed) {
    if (timed) {
        for (int i = 0; i < 2; i++)
            if (!f.ev[i]) BCHK(hipEventCreate(&f.ev[i]));
    }
    if (!b->bc.use_graphs) {
        if (timed) BCHK(hipEventRecord(f.ev[0], b->stream));
        for (int s = 0; s < n; s++) issue_step(b, f);
        BCHK(hipGetLastError());
        if (timed) BCHK(hipEventRecord(f.ev[1], b->stream));
        return TS_OK;
    }
    if (!f.graph || f.graph_steps != n) {
        if (f.graph) BCHK(hipGraphExecDestroy(f.graph));
        f.graph = nullptr;
        hipGraph_t gr;
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < n; s++) issue_step(b, f);
        BCHK(hipStreamEndCapture(b->stream, &gr));
        hipError_t e = hipGraphInstantiate(&f.graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            f.graph = nullptr;
            return fail(b, TS_ERR_HIP, tsr::hip_msg(e, "hipGraphInstantiate"));
        }
        f.graph_steps = n;
    }
    if (timed) BCHK(hipEventRecord(f.ev[0], b->stream));
    BCHK(hipGraphLaunch(f.graph, b->stream));
    if (timed) BCHK(hipEventRecord(f.ev[1], b->stream));
    return TS_OK;
}

// after the pass has synchronised: credit the timed pass to its candidate, and pick the bucket's
// winner once every candidate has its passes (a new candidate or the winner needs a retile)
int credit_trial(ts_batch* b, Family& f) {
    if (!f.timed) return TS_OK;
    f.timed = false;
    float ms = 0;
    BCHK(hipEventElapsedTime(&ms, f.ev[0], f.ev[1]));
    LoadTuning& lt = f.tuning[f.bucket];
    if (lt.chosen >= 0 || f.cand < 0) return TS_OK;
    TileTrial& t = lt.t[f.cand];
    t.ms += ms;
    t.cell_steps += f.pass_cell_steps;
    t.passes++;
    bool all = true;
    for (const TileTrial& x : lt.t) all = all && x.passes >= kTrialPasses;
    if (all) {
        int k = 0;
        for (int c = 1; c < kTileCands; c++)
            if (lt.t[c].ms * lt.t[k].cell_steps < lt.t[k].ms * lt.t[c].cell_steps) k = c;
        lt.chosen = k;
    }
    if (next_candidate(lt) != f.cand) f.dirty = true;
    return TS_OK;
}

// Issue, on the copy stream, the conversion of a finished problem's state to the reference AoS
// layout (ghosts zero) in its region of the outgoing staging buffer and the copy to the host.
// Needs no lock: the problem is out of its family but not yet published, so poll does not read it.
int issue_retire(ts_batch* b, Problem& p, size_t region) {
    Layout L = layout_of(p.cfg);
    const size_t na = tsr::aos_doubles(p.cfg.ndims, p.n);
    double* st = b->d_rstage + region;
    p.result.resize(na);
    BCHK(hipMemsetAsync(st, 0, na * sizeof(double), b->cstream));
    BCHK(tsr::device_soa_to_aos((const double*)(b->d_arena + p.off + L.u0), st, p.cfg.ndims, p.n, b->cstream));
    BCHK(hipMemcpyAsync(p.result.data(), st, na * sizeof(double), hipMemcpyDeviceToHost, b->cstream));
    return TS_OK;
}

// control fields -> step info, marks the problem done and frees its block (under the lock)
void finish_retire(ts_batch* b, Problem& p, const StepCtl& ctl) {
    p.info.time = ctl.time;
    p.info.dt = ctl.dt;
    p.info.iteration = ctl.iter;
    p.info.steps_done = ctl.steps_done;
    for (int v = 0; v < 5; v++) p.info.residual_norm[v] = 0;
    p.error = ctl.error;
    p.status = 2;
    b->arena.release(p.off, p.bytes);
}

}  // namespace

extern "C" {

int ts_batch_create(const ts_batch_config* bcfg, ts_batch** out) {
    if (!out) return TS_ERR_INVALID;
    *out = nullptr;
    ts_batch* b = new ts_batch();
    b->families.reserve(kMaxFamilies);
    if (bcfg) b->bc = *bcfg;
    if (b->bc.hbm_budget_bytes == 0) b->bc.hbm_budget_bytes = (size_t)8 << 30;
    if (b->bc.max_active <= 0) b->bc.max_active = 1 << 20;
    hipError_t e;
    if ((e = hipSetDevice(b->bc.device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_arena, b->bc.hbm_budget_bytes)) != hipSuccess) {
        tsr::set_thread_error("ts_batch_create: " + tsr::hip_msg(e, "device set-up"));
        if (b->stream) (void)hipStreamDestroy(b->stream);
        if (b->cstream) (void)hipStreamDestroy(b->cstream);
        delete b;
        return e == hipErrorOutOfMemory ? TS_ERR_NOMEM : TS_ERR_HIP;
    }
    b->arena.init(b->d_arena, b->bc.hbm_budget_bytes);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->bc.device) == hipSuccess) b->ncu = prop.multiProcessorCount;
    *out = b;
    return TS_OK;
}

int ts_batch_destroy(ts_batch* b) {
    if (!b) return TS_ERR_INVALID;
    ts_batch_stop_server(b);
    (void)hipSetDevice(b->bc.device);
    (void)hipStreamSynchronize(b->stream);
    if (b->cstream) (void)hipStreamSynchronize(b->cstream);
    for (Family& f : b->families) {
        if (f.graph) (void)hipGraphExecDestroy(f.graph);
        if (f.d_desc) (void)hipFree(f.d_desc);
        if (f.d_table) (void)hipFree(f.d_table);
        if (f.d_ntiles) (void)hipFree(f.d_ntiles);
        if (f.d_part) (void)hipFree(f.d_part);
        if (f.d_ranges) (void)hipFree(f.d_ranges);
        if (f.d_nmem) (void)hipFree(f.d_nmem);
        if (f.d_fresh) (void)hipFree(f.d_fresh);
        if (f.d_fresh_idx) (void)hipFree(f.d_fresh_idx);
        for (hipEvent_t e : f.ev)
            if (e) (void)hipEventDestroy(e);
    }
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->h_ptr) (void)hipHostFree(b->h_ptr);
    if (b->d_ctl) (void)hipFree(b->d_ctl);
    if (b->d_ptr) (void)hipFree(b->d_ptr);
    if (b->d_arena) (void)hipFree(b->d_arena);
    if (b->d_stage) (void)hipFree(b->d_stage);
    if (b->d_rstage) (void)hipFree(b->d_rstage);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->cstream) (void)hipStreamDestroy(b->cstream);
    delete b;
    return TS_OK;
}

const char* ts_batch_last_error(const ts_batch* b) {
    if (!b) return "null batch";
    // a copy per calling thread: the server thread may record a new error meanwhile
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(const_cast<ts_batch*>(b)->err_mu);
    copy = b->err;
    return copy.c_str();
}

int ts_batch_submit(ts_batch* b, const ts_config* cfg, const double* init_state, int max_steps,
                    double t_end, long long* ticket) {
    if (!b || !cfg || !init_state || !ticket || max_steps < 0) return TS_ERR_INVALID;
    // the same validation as ts_create (cell-count limit of the 32-bit plane offsets, known
    // precision, Riemann solver and boundary types, ...), then the batch's own limits
    {
        std::string err;
        int rc = tsr::validate(cfg, err);
        if (rc != TS_OK) return fail(b, rc, err);
    }
    if (cfg->ndims < 1 || cfg->ndims > 2) return fail(b, TS_ERR_UNSUPPORTED, "batched solves are 1-D or 2-D");
    if (cfg->rk_stages != 3) return fail(b, TS_ERR_UNSUPPORTED, "batched solves use SSP-RK3");
    if (cfg->viscous || cfg->external_force || cfg->external_accel || cfg->viscosity != 0.0)
        return fail(b, TS_ERR_UNSUPPORTED, "viscous terms and body forces are not batched");
    if (cfg->ghost != 1 || cfg->nx < 2 || (cfg->ndims > 1 && cfg->ny < 2) || cfg->nz != 1 ||
        (cfg->ndims == 1 && cfg->ny != 1) || !(cfg->gamma > 1.0) || !(cfg->cfl > 0.0) || cfg->n_bc < 1 ||
        cfg->n_bc > TS_MAX_BC)
        return fail(b, TS_ERR_INVALID, "invalid configuration");
    for (int s = 0; s < 6; s++)
        if (cfg->face_marker[s]) return fail(b, TS_ERR_UNSUPPORTED, "per-face markers are not batched");
    for (int s = 0; s < 2 * cfg->ndims; s++)
        if (!cfg->periodic[s / 2] && (cfg->side_marker[s] < 0 || cfg->side_marker[s] >= cfg->n_bc))
            return fail(b, TS_ERR_INVALID, "side marker refers to a missing boundary condition");
    auto p = std::make_unique<Problem>();
    p->cfg = *cfg;
    // the caller's grid arrays may be gone by admission time: keep copies
    for (int a = 0; a < 3; a++) {
        const int na = (a == 0 ? cfg->nx : a == 1 ? cfg->ny : cfg->nz) + 2;
        p->cfg.coord[a] = p->cfg.h[a] = nullptr;
        if (a < cfg->ndims && cfg->coord[a]) {
            p->coord[a].assign(cfg->coord[a], cfg->coord[a] + na);
            p->cfg.coord[a] = p->coord[a].data();
        }
        if (a < cfg->ndims && cfg->h[a]) {
            p->h[a].assign(cfg->h[a], cfg->h[a] + na);
            p->cfg.h[a] = p->h[a].data();
        }
    }
    p->n[0] = cfg->nx;
    p->n[1] = cfg->ny;
    p->n[2] = cfg->nz;
    p->ncell = state_cells(*cfg);
    p->max_steps = max_steps;
    p->t_end = t_end;
    p->bytes = layout_of(*cfg).total;
    if (p->bytes > b->arena.capacity) return fail(b, TS_ERR_NOMEM, "solve is larger than the HBM budget");
    size_t nst = ts_state_size_cfg(cfg);
    p->init.assign(init_state, init_state + nst);
    std::lock_guard<std::mutex> lk(b->mu);
    int fam = family_of(b, *cfg);
    if (fam < 0) return fail(b, TS_ERR_UNSUPPORTED, "no batched kernel for this configuration");
    p->family = fam;
    p->ticket = b->next_ticket++;
    *ticket = p->ticket;
    b->queue.push_back(p->ticket);
    b->probs[p->ticket] = std::move(p);
    b->cv.notify_all();
    return TS_OK;
}

// wait for the outgoing copies in flight and publish those solves (done, memory released)
static int complete_retiring(ts_batch* b) {
    if (b->retiring.empty()) return TS_OK;
    BCHK(hipStreamSynchronize(b->cstream));
    std::lock_guard<std::mutex> lk(b->mu);
    for (Problem* p : b->retiring) {
        std::vector<double>().swap(p->init);
        finish_retire(b, *p, p->final_ctl);
    }
    b->retiring.clear();
    return TS_OK;
}

// Admission: take queued solves off the queue in arrival order while they fit the budget,
// upload them (copy stream) and attach them to their families.
static int admit_queued(ts_batch* b) {
    // (1) under the lock: reserve their arena blocks (head-of-line waits for memory)
    std::vector<Problem*> incoming;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        size_t nres = 0;
        for (Family& f : b->families) nres += f.members.size();
        while (!b->queue.empty() && (int)(nres + incoming.size()) < b->bc.max_active) {
            Problem& p = *b->probs[b->queue.front()];
            size_t off;
            if (!b->arena.alloc(p.bytes, &off)) break;
            p.off = off;
            p.status = 1;
            b->queue.pop_front();
            incoming.push_back(&p);
        }
    }
    if (incoming.empty()) return TS_OK;
    // (2) without the lock: upload their geometry, state and control blocks, so that submit and
    // poll on other threads do not wait behind the copies
    std::vector<int> admitted(incoming.size(), 0);
    for (size_t i = 0; i < incoming.size(); i++) admitted[i] = admit(b, *incoming[i]) == TS_OK;
    BCHK(hipStreamSynchronize(b->cstream));
    for (Problem* p : incoming) {
        std::vector<double>().swap(p->init);
        std::vector<double>().swap(p->geo_host);
    }
    // (3) under the lock again: attach them to their families (or fail them)
    std::lock_guard<std::mutex> lk(b->mu);
    for (size_t i = 0; i < incoming.size(); i++) {
        Problem& p = *incoming[i];
        if (!admitted[i]) {
            b->arena.release(p.off, p.bytes);
            p.status = 2;
            p.error = 1;
            continue;
        }
        Family& fam = b->families[p.family];
        if (fam.free_slots.empty()) {
            p.slot = fam.slot_end++;
        } else {
            p.slot = fam.free_slots.back();
            fam.free_slots.pop_back();
        }
        p.uploaded = false;
        fam.members.push_back(&p);
        b->families[p.family].dirty = true;
    }
    return TS_OK;
}

// one admission + advance pass; the caller is either the server thread or, when no server
// runs, the user through ts_batch_run.  Runs under run_mu.  While solves are resident, the
// queued ones are uploaded after this pass's kernels are issued (the copies overlap them) and
// take their first steps in the next pass; an idle batch admits them before its kernels.
static int batch_run(ts_batch* b, int n_steps, int* pending) {
    BCHK(hipSetDevice(b->bc.device));
    bool busy;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        size_t nres = 0;
        for (Family& f : b->families) nres += f.members.size();
        busy = nres > 0 && n_steps > 0;
    }
    if (!busy) {
        int r = admit_queued(b);
        if (r) return r;
    }
    std::vector<Problem*> resident;
    size_t nfam = 0;  // families that exist now; later ones are appended without moving these
    {
        std::lock_guard<std::mutex> lk(b->mu);
        nfam = b->families.size();
    }
    for (size_t fi = 0; fi < nfam; fi++)
        for (Problem* q : b->families[fi].members) resident.push_back(q);
    if (!resident.empty() && n_steps > 0) {
        for (size_t fi = 0; fi < nfam; fi++) {
            Family& f = b->families[fi];
            if (f.members.empty()) continue;
            if (f.dirty) {
                int r = rebuild_tables(b, f);
                if (r) return r;
            }
            const bool timed = f.cand >= 0 && f.tuning[f.bucket].chosen < 0;
            int r = issue_steps(b, f, n_steps, timed);
            if (r) return r;
            if (timed) {
                f.timed = true;
                f.pass_cell_steps = 0;
            }
        }
        // the previous pass's finished solves, copied out while these kernels run (their arena
        // blocks are free after this), then the queued solves' uploads, also under the kernels
        {
            int r = complete_retiring(b);
            if (r) return r;
            if (busy && (r = admit_queued(b))) return r;
        }
        // control blocks of every resident solve back to the host
        if (b->h_ctl_cap < resident.size()) {
            BCHK(hipStreamSynchronize(b->stream));
            if (b->h_ctl) BCHK(hipHostFree(b->h_ctl));
            if (b->h_ptr) BCHK(hipHostFree(b->h_ptr));
            if (b->d_ctl) BCHK(hipFree(b->d_ctl));
            if (b->d_ptr) BCHK(hipFree(b->d_ptr));
            b->h_ctl = nullptr, b->h_ptr = nullptr, b->d_ctl = nullptr, b->d_ptr = nullptr;
            b->h_ctl_cap = 0;
            const size_t cap = resident.size() * 2;
            BCHK(hipHostMalloc((void**)&b->h_ctl, cap * sizeof(StepCtl), hipHostMallocDefault));
            BCHK(hipHostMalloc((void**)&b->h_ptr, cap * sizeof(StepCtl*), hipHostMallocDefault));
            BCHK(hipMalloc((void**)&b->d_ctl, cap * sizeof(StepCtl)));
            BCHK(hipMalloc((void**)&b->d_ptr, cap * sizeof(StepCtl*)));
            b->h_ctl_cap = cap;
        }
        const int nres = (int)resident.size();
        for (int i = 0; i < nres; i++) b->h_ptr[i] = resident[i]->desc.ctl;
        BCHK(hipMemcpyAsync(b->d_ptr, b->h_ptr, nres * sizeof(StepCtl*), hipMemcpyHostToDevice, b->stream));
        const int words = nres * int(sizeof(StepCtl) / 8);
        hipLaunchKernelGGL(gather_ctl, dim3((words + 255) / 256), dim3(256), 0, b->stream, b->d_ptr, b->d_ctl, nres);
        BCHK(hipGetLastError());
        BCHK(hipMemcpyAsync(b->h_ctl, b->d_ctl, nres * sizeof(StepCtl), hipMemcpyDeviceToHost, b->stream));
        BCHK(hipStreamSynchronize(b->stream));
        // the steps each solve advanced in this pass (a solve that stops part-way does fewer than
        // n_steps), summed per family for the autotuner
        for (int i = 0; i < nres; i++) {
            Problem& p = *resident[i];
            const int d = b->h_ctl[i].steps_done - p.steps_seen;
            p.steps_seen = b->h_ctl[i].steps_done;
            Family& f = b->families[p.family];
            if (f.timed) f.pass_cell_steps += double(p.ncell) * d;
        }
        for (size_t fi = 0; fi < nfam; fi++) {
            int r = credit_trial(b, b->families[fi]);
            if (r) return r;
        }
        // finished solves: their states go out on the copy stream (the kernels that wrote them
        // have completed: the stream was synchronised above), one staging region each
        std::vector<int> fin;
        size_t need = 0;
        for (int i = 0; i < nres; i++) {
            const StepCtl& c = b->h_ctl[i];
            if (c.active && !c.error) continue;
            fin.push_back(i);
            need += tsr::aos_doubles(resident[i]->cfg.ndims, resident[i]->n);
        }
        if (!fin.empty()) {
            if (need > b->rstage_cap) {  // the copy stream is idle: complete_retiring ran above
                if (b->d_rstage) BCHK(hipFree(b->d_rstage));
                b->d_rstage = nullptr;
                b->rstage_cap = 0;
                BCHK(hipMalloc((void**)&b->d_rstage, need * 2 * sizeof(double)));
                b->rstage_cap = need * 2;
            }
            size_t region = 0;
            for (int i : fin) {
                Problem& p = *resident[i];
                p.final_ctl = b->h_ctl[i];
                int r = issue_retire(b, p, region);
                if (r) return r;
                region += tsr::aos_doubles(p.cfg.ndims, p.n);
            }
            std::lock_guard<std::mutex> lk(b->mu);
            for (int i : fin) {
                Problem& p = *resident[i];
                Family& f = b->families[p.family];
                f.members.erase(std::find(f.members.begin(), f.members.end(), &p));
                f.free_slots.push_back(p.slot);
                f.dirty = true;
                b->retiring.push_back(&p);
            }
        }
    }
    // with no solve left to advance, nothing would overlap the outgoing copies: publish them now
    {
        bool idle;
        {
            std::lock_guard<std::mutex> lk(b->mu);
            size_t nres = 0;
            for (Family& f : b->families) nres += f.members.size();
            idle = nres == 0 || n_steps <= 0;
        }
        if (idle) {
            int r = complete_retiring(b);
            if (r) return r;
        }
    }
    if (pending) {
        std::lock_guard<std::mutex> lk(b->mu);
        size_t n = b->queue.size() + b->retiring.size();
        for (Family& f : b->families) n += f.members.size();
        *pending = (int)n;
    }
    b->cv.notify_all();
    return TS_OK;
}

int ts_batch_run(ts_batch* b, int n_steps, int* pending) {
    if (!b || n_steps < 0) return TS_ERR_INVALID;
    if (b->serving) return fail(b, TS_ERR_INVALID, "the server thread advances this batch: submit and poll instead");
    // a server started meanwhile waits for this pass: passes never overlap
    std::lock_guard<std::mutex> run(b->run_mu);
    return batch_run(b, n_steps, pending);
}

int ts_batch_poll(ts_batch* b, long long ticket, int* done, double* out_state, ts_step_info* info) {
    if (!b || !done) return TS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->probs.find(ticket);
    if (it == b->probs.end()) return fail(b, TS_ERR_INVALID, "unknown ticket");
    Problem& p = *it->second;
    if (p.status != 2) {
        *done = 0;
        return TS_OK;
    }
    *done = 1;
    if (out_state) std::memcpy(out_state, p.result.data(), p.result.size() * sizeof(double));
    if (info) *info = p.info;
    int err = p.error;
    b->probs.erase(it);
    return err ? TS_ERR_UNPHYSICAL : TS_OK;
}

int ts_batch_start_server(ts_batch* b) {
    if (!b) return TS_ERR_INVALID;
    std::lock_guard<std::mutex> g(b->server_mu);
    if (b->serving) return TS_OK;
    b->stop = false;
    b->serving = true;
    b->server = std::thread([b] {
        while (!b->stop) {
            {
                std::unique_lock<std::mutex> lk(b->mu);
                bool work = !b->queue.empty();
                for (Family& f : b->families) work = work || !f.members.empty();
                if (!work) {
                    b->cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
            int pending = 0, rc;
            {
                std::lock_guard<std::mutex> run(b->run_mu);
                rc = batch_run(b, b->steps_per_chunk, &pending);
            }
            if (rc != TS_OK) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    });
    return TS_OK;
}

int ts_batch_stop_server(ts_batch* b) {
    if (!b) return TS_ERR_INVALID;
    std::lock_guard<std::mutex> g(b->server_mu);
    if (!b->serving) return TS_OK;
    // the server finishes its pass and exits before ts_batch_run is allowed again
    b->stop = true;
    b->cv.notify_all();
    if (b->server.joinable()) b->server.join();
    b->serving = false;
    return TS_OK;
}

int ts_batch_tuning(ts_batch* b, ts_tile_trial* out, int cap, int* n) {
    if (!b || !n || cap < 0 || (cap > 0 && !out)) return TS_ERR_INVALID;
    std::lock_guard<std::mutex> run(b->run_mu);  // the pass that updates the trials is not running
    size_t nfam;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        nfam = b->families.size();
    }
    int k = 0;
    for (size_t fi = 0; fi < nfam; fi++) {
        const Family& f = b->families[fi];
        for (const auto& kv : f.tuning)
            for (int c = 0; c < kTileCands; c++) {
                const TileTrial& t = kv.second.t[c];
                if (t.passes == 0) continue;
                if (k < cap) {
                    ts_tile_trial& o = out[k];
                    o.ndims = std::get<0>(f.key);
                    o.riemann = std::get<1>(f.key);
                    o.recon_variant = std::get<2>(f.key);
                    o.precision = std::get<3>(f.key);
                    o.load_bucket = kv.first;
                    o.tile_param = o.ndims == 1 ? kCand1D[c] : kCand2D[c];
                    o.passes = t.passes;
                    o.ns_per_cell_step = t.cell_steps > 0 ? t.ms * 1e6 / t.cell_steps : 0.0;
                    o.chosen = kv.second.chosen == c ? 1 : 0;
                }
                k++;
            }
    }
    *n = k < cap ? k : cap;
    return TS_OK;
}

}  // extern "C"
