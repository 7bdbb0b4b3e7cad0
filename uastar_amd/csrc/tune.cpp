// tune.cpp -- the autotuner: picks, per instance class, the team width (waves = heaps
// per query) of batched and served queries, the hash load factor of puzzle queries and
// the heap count of the plain whole-GPU solve, by running a caller-given sample of
// queries under each candidate and measuring it on the device.
//
// uAstar fixes all of these at compile time (39 blocks x 192 threads, one heap per
// thread, 1,335 / 133 entries per heap, a 90,000,089-entry puzzle table;
// pathway/GPU-solver.hpp:5-17, puzzle/GPU-solver.cuh:12-22).
//
// Cost model for a team width w (blocks per team b = w / 4): the sample of one class is
// solved in one batch, each query on its own team, giving its team time t_i (device
// clock stamps of the team). A stream of such queries on B resident blocks of the class
// finishes no sooner than its total team-block time over B and no sooner than its
// longest query, so the estimate is max(sum_i b * t_i / B, max_i t_i). The candidate
// with the lowest estimate wins; a wider team must beat a narrower one by 3 %.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace uart;

namespace {

struct Trial {
    double est_ms = 0;    // the stream-time estimate above
    double max_ins = 0;   // most hash insertions of one query (puzzles)
    bool ok = false;
};

Trial measure_batch(ua_ctx *c, const std::vector<ua_query> &qs, int bpg, int class_blocks, int *rc) {
    Trial t;
    std::vector<ua_result> res(qs.size());
    *rc = ua_batch_solve(c, qs.data(), (int64_t)qs.size(), res.data());
    if (*rc < 0) return t;
    double sum = 0, mx = 0, ins = 0;
    for (const ua_result &r : res) {
        if (r.status < 0) return t;  // a candidate that cannot run the sample loses
        sum += r.device_ms;
        mx = std::max(mx, r.device_ms);
        ins = std::max(ins, (double)r.inserted);
    }
    t.est_ms = std::max(sum * bpg / std::max(1, class_blocks), mx);
    t.max_ins = ins;
    t.ok = true;
    return t;
}

// log2 of the hash slots that keep the sample's largest query at load <= 1/2, plus one
// doubling of headroom for instances harder than the sample.
int hash_for(double max_ins, int lo, int hi) {
    double need = std::max(1.0, max_ins) * 2.0;
    int hl = (int)std::ceil(std::log2(need)) + 1;
    return std::min(hi, std::max(lo, hl));
}

}  // namespace

extern "C" int ua_get_tuning(ua_ctx *c, ua_tuning *o) {
    if (!c || !o) return fail(c, UA_ERR_INVALID, "null argument");
    std::memset(o, 0, sizeof *o);
    o->grid_team_waves = 4 * team_bpg(c, KIND_GRID);
    o->puzzle_team_waves = 4 * team_bpg(c, KIND_PUZZLE);
    o->puzzle_hash_log2 = batch_hash_log2(c, 4);
    o->plain_grid_heaps = plain_heaps(c, KIND_GRID);
    o->plain_puzzle_heaps = plain_heaps(c, KIND_PUZZLE);
    o->plain_hash_log2 = plain_hash_log2(c, 4);
    return UA_OK;
}

extern "C" int ua_set_tuning(ua_ctx *c, const ua_tuning *t) {
    if (!c || !t) return fail(c, UA_ERR_INVALID, "null argument");
    auto in = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
    if (!in(t->grid_team_waves, 4, 256) || !in(t->puzzle_team_waves, 4, 256) || !in(t->puzzle_hash_log2, 10, 32) ||
        !in(t->plain_grid_heaps, 4, 1 << 20) || !in(t->plain_puzzle_heaps, 4, 1 << 20) ||
        !in(t->plain_hash_log2, 10, 32))
        return fail(c, UA_ERR_INVALID, "tuning value out of range");
    c->tune.grid_waves = t->grid_team_waves;
    c->tune.puzzle_waves = t->puzzle_team_waves;
    c->tune.puzzle_hash_log2[4] = t->puzzle_hash_log2;
    c->tune.plain_heaps[0] = t->plain_grid_heaps;
    c->tune.plain_heaps[1] = t->plain_puzzle_heaps;
    c->tune.plain_hash_log2[4] = t->plain_hash_log2;
    c->tune.tuned = true;
    return UA_OK;
}

extern "C" int ua_autotune(ua_ctx *c, const ua_query *qs, int64_t n, ua_tuning *out) {
    if (!c || !qs || n <= 0 || !out) return fail(c, UA_ERR_INVALID, "bad arguments");
    std::vector<ua_query> grids, puzzles;
    for (int64_t i = 0; i < n; ++i) {
        if (qs[i].kind == UA_QUERY_GRID) grids.push_back(qs[i]);
        else if (qs[i].kind == UA_QUERY_PUZZLE && qs[i].n == 4) puzzles.push_back(qs[i]);
        else if (qs[i].kind != UA_QUERY_PUZZLE) return fail(c, UA_ERR_INVALID, "sample query %lld: bad kind", (long long)i);
    }
    int per_cu = 0;
    HIPCHK(c, teams_max_blocks(&per_cu));
    const int resident = std::max(1, c->num_cu * std::max(1, per_cu));
    const bool both = !grids.empty() && !puzzles.empty();
    const int blocks_puzzle = both ? resident / 2 : resident;
    const int blocks_grid = both ? resident - resident / 2 : resident;
    int trials = 0, rc = 0;
    double est_grid = 0, est_puzzle = 0;
    const Tuning saved = c->tune;
    auto bail = [&](int code) {
        c->tune = saved;
        return code;
    };

    // 1. grid team width
    if (!grids.empty() && c->cfg.team_waves <= 0) {
        int best_w = c->tune.grid_waves;
        double best = 0;
        for (int w : {4, 8, 16}) {
            c->tune.grid_waves = w;
            Trial t = measure_batch(c, grids, w / 4, blocks_grid, &rc);
            if (rc < 0) return bail(rc);
            ++trials;
            if (t.ok && (best == 0 || t.est_ms < 0.97 * best)) {
                best = t.est_ms;
                best_w = w;
            }
        }
        c->tune.grid_waves = best_w;
        est_grid = best;
    }
    // 2. puzzle team width, then the hash size from the chosen width's insertions
    if (!puzzles.empty() && c->cfg.puzzle_team_waves <= 0) {
        int best_w = c->tune.puzzle_waves;
        double best = 0, best_ins = 0;
        for (int w : {16, 32, 64, 128}) {
            c->tune.puzzle_waves = w;
            Trial t = measure_batch(c, puzzles, w / 4, blocks_puzzle, &rc);
            if (rc < 0) return bail(rc);
            ++trials;
            if (t.ok && (best == 0 || t.est_ms < 0.97 * best)) {
                best = t.est_ms;
                best_w = w;
                best_ins = t.max_ins;
            }
        }
        c->tune.puzzle_waves = best_w;
        est_puzzle = best;
        if (best > 0 && c->cfg.hash_log2 <= 0) c->tune.puzzle_hash_log2[4] = hash_for(best_ins, 18, 28);
    }
    // 3. heaps of the plain whole-GPU solve, per kind, on the first few sample queries
    if (c->cfg.plain_heaps <= 0) {
        const int max_heaps = 4 * resident;
        for (int kind : {KIND_GRID, KIND_PUZZLE}) {
            const std::vector<ua_query> &src = kind == KIND_GRID ? grids : puzzles;
            if (src.empty()) continue;
            const size_t m = std::min<size_t>(3, src.size());
            int best_h = 0;
            double best = 0, best_ins = 0;
            for (int h = 256; h <= std::min(4096, max_heaps); h *= 2) {
                c->tune.plain_heaps[kind == KIND_GRID ? 0 : 1] = h;
                double ms = 0, ins = 0;
                bool ok = true;
                for (size_t k = 0; k < m && ok; ++k) {
                    ua_result r;
                    const ua_query &u = src[k];
                    int s = kind == KIND_GRID ? ua_grid_solve(c, u.grid, u.sx, u.sy, u.ex, u.ey, &r)
                                              : ua_puzzle_solve(c, u.n, u.tiles, &r);
                    if (s < 0 && s != UA_ERR_CAPACITY && s != UA_ERR_ROUNDS) return bail(s);
                    ok = s >= 0;
                    ms += r.kernel_ms;
                    ins = std::max(ins, (double)r.inserted);
                }
                ++trials;
                if (ok && (best == 0 || ms < 0.97 * best)) {
                    best = ms;
                    best_h = h;
                    best_ins = ins;
                }
            }
            c->tune.plain_heaps[kind == KIND_GRID ? 0 : 1] = best_h;
            if (kind == KIND_PUZZLE && best_h && c->cfg.hash_log2 <= 0)
                c->tune.plain_hash_log2[4] = hash_for(best_ins, 18, 28);
        }
    }
    c->tune.tuned = true;
    rc = ua_get_tuning(c, out);
    out->trials = trials;
    out->est_grid_ms = est_grid;
    out->est_puzzle_ms = est_puzzle;
    return rc;
}
