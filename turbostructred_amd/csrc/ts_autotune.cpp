// ts_autotune.cpp -- the tile-shape autotuner and its process-wide table.
//
// The only free tiling parameter of the marching kernels is the number of output rows per tile
// (cells per thread along the march).  Results never depend on it: every cell's arithmetic is
// the same whatever the tiling, so the table may be keyed by problem size alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ts_internal.hpp"

namespace {
using Key = std::tuple<int, int, int, int, int, int>;  // nd, riemann, recon variant, precision, stages, size bucket
std::mutex g_mu;
std::map<Key, int> g_table;

int size_bucket(const ts_config& c) {
    long long n = (long long)c.nx * c.ny * c.nz;
    int b = 0;
    while (n > 1) { n >>= 1; b++; }
    return b;
}

Key key_of(const ts_config& c) {
    return Key(c.ndims, c.riemann, tsr::recon_variant(c), c.precision, c.rk_stages, size_bucket(c));
}
}  // namespace

namespace tsr {
int tuned_rows_per_tile(const ts_config& c, long long slots) {
    if (c.ndims == 1) return 1;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_table.find(key_of(c));
        if (it != g_table.end()) return it->second;
    }
    const long long strips = (c.nx + 59) / 60;
    const long long zb = c.ndims == 3 ? (c.nz + 5) / 6 : 1;  // z blocks of a 3-D tile column
    if (slots <= 0) {
        // no occupancy known: enough tiles to fill 256 CUs several times over, at most 32 rows
        long long rows = (long long)c.ny * zb;
        int ty = 32;
        while (ty > 4 && strips * ((rows + ty - 1) / ty) < 4096) ty /= 2;
        return ty;
    }
    // A tile is one wavefront's (3-D: one workgroup's) march of ty rows plus two warm-up rows,
    // and tiles run in rounds of `slots`.  The time goes as rounds * (ty + 2): pick the ty that
    // minimises it, so that the last round is not a mostly empty one.
    int best = 4;
    double best_cost = 1e300;
    const int hi = c.ny < 256 ? c.ny : 256;
    for (int ty = 4; ty <= (hi > 4 ? hi : 4); ty++) {
        long long tiles = strips * zb * ((c.ny + ty - 1) / ty);
        long long rounds = (tiles + slots - 1) / slots;
        double cost = double(rounds) * (ty + 2);
        if (cost < best_cost) {
            best_cost = cost;
            best = ty;
        }
    }
    return best;
}
}  // namespace tsr

namespace {
// Time `steps` steps of cfg from the reference AoS state (NULL: a quiescent uniform state) for
// each candidate tile height; the winner goes into the table.
int autotune_impl(const ts_config* cfg, const double* aos, size_t n_doubles, int steps, int* cand_rows,
                  double* cand_ms, int cap, int* n_cand, int* chosen, double* best_ms) {
    if (n_cand) *n_cand = 0;
    if (!cfg || steps < 1 || (aos && n_doubles != ts_state_size_cfg(cfg))) return TS_ERR_INVALID;
    if (cfg->ndims == 1) {
        if (chosen) *chosen = 1;
        if (best_ms) *best_ms = 0;
        return TS_OK;
    }
    // the occupancy model's choice (with no table entry) and its neighbours, the heights at which
    // the tiles just fit one and two rounds of the resident slots, then fixed heights
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_table.erase(key_of(*cfg));
    }
    std::vector<int> cands;
    {
        ts_ctx* c = nullptr;
        if (ts_create(cfg, &c) == TS_OK) {
            const int m = c->ty;
            cands = {m, m > 5 ? m - m / 4 : 4, m + m / 4 + 1};
            const long long strips = c->nstrips, tiles_m = c->ntiles;
            const long long zb = cfg->ndims == 3 ? (cfg->nz + tsd::ZOUT - 1) / tsd::ZOUT : 1;
            // slots implied by the model's choice: its tile count rounded up to a whole round
            for (long long rounds = 1; rounds <= 2; rounds++) {
                const long long target = tiles_m * rounds;
                for (int ty = 4; ty <= 256; ty++)
                    if (strips * zb * ((cfg->ny + ty - 1) / ty) <= target) {
                        cands.push_back(ty);
                        break;
                    }
            }
            ts_destroy(c);
        }
    }
    for (int ty : {4, 8, 16, 24, 32, 48, 64}) cands.push_back(ty);
    const int ymax = cfg->ny;
    std::vector<int> uniq;
    for (int ty : cands)
        if (ty >= 1 && ty <= std::max(ymax, 4) && std::find(uniq.begin(), uniq.end(), ty) == uniq.end()) uniq.push_back(ty);
    const size_t nst = ts_state_size_cfg(cfg);
    std::vector<double> quiet;
    if (!aos) {
        // a quiescent uniform state: the arithmetic of the flux kernels barely depends on the data
        quiet.assign(nst, 0.0);
        for (size_t i = 0; i < nst / 5; i++) {
            quiet[i * 5 + 0] = 1.0;
            quiet[i * 5 + 4] = 2.5;
        }
        aos = quiet.data();
    }
    int best = -1, k = 0;
    double bms = 1e300;
    for (int ty : uniq) {
        {
            std::lock_guard<std::mutex> lk(g_mu);
            g_table[key_of(*cfg)] = ty;
        }
        ts_ctx* c = nullptr;
        if (ts_create(cfg, &c) != TS_OK) continue;
        ts_step_info info;
        int rc = ts_set_state(c, aos, nst);
        if (rc == TS_OK) rc = ts_step(c, 2, 1e300, 0, &info);  // warm-up
        double ms = 0;
        long long n = 0;
        if (rc == TS_OK) {
            ts_enable_kernel_timing(c, 1);
            rc = ts_step(c, steps, 1e300, 0, &info);
            ts_get_kernel_timing(c, 4, &ms, &n);
            if (info.steps_done < steps) rc = TS_ERR_UNPHYSICAL;  // the state ended the run early
        }
        ts_destroy(c);
        if (rc != TS_OK || n == 0) continue;
        ms /= steps;
        if (cand_rows && cand_ms && k < cap) {
            cand_rows[k] = ty;
            cand_ms[k] = ms;
        }
        k++;
        if (ms < bms) {
            bms = ms;
            best = ty;
        }
    }
    if (n_cand) *n_cand = std::min(k, cap);
    std::lock_guard<std::mutex> lk(g_mu);
    if (best < 0) {
        g_table.erase(key_of(*cfg));
        return TS_ERR_UNSUPPORTED;
    }
    g_table[key_of(*cfg)] = best;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = bms;
    return TS_OK;
}
}  // namespace

extern "C" int ts_autotune(const ts_config* cfg, int* chosen, double* best_ms) {
    return autotune_impl(cfg, nullptr, 0, 5, nullptr, nullptr, 0, nullptr, chosen, best_ms);
}

extern "C" int ts_autotune_state(const ts_config* cfg, const double* aos_state, size_t n_doubles, int steps,
                                 int* cand_rows, double* cand_ms_per_step, int cap, int* n_cand, int* chosen) {
    if (cap < 0 || (cap > 0 && (!cand_rows || !cand_ms_per_step))) return TS_ERR_INVALID;
    return autotune_impl(cfg, aos_state, n_doubles, steps, cand_rows, cand_ms_per_step, cap, n_cand, chosen, nullptr);
}
