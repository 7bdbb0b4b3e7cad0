// tune.cpp -- the autotuner: picks, per instance class (15-puzzles, grids below and
// at or above a size threshold), the team width of batched and served queries (whole
// blocks of 16 wavefronts: 16, 32, 64 or 128 waves), the hash load factor of puzzle
// queries and the heap count of the plain whole-GPU solve, by running a caller-given
// sample of queries under each candidate and measuring it on the device.
//
// uAstar fixes all of these at compile time (39 blocks x 192 threads, one heap per
// thread, 1,335 / 133 entries per heap, a 90,000,089-entry puzzle table;
// pathway/GPU-solver.hpp:5-17, puzzle/GPU-solver.cuh:12-22).
//
// Cost model for a team width w: the sample of one class is solved in one batch, each
// query on its own team, giving its team time t_i (device clock stamps of the team).
// A stream of such queries on S resident wave slots finishes no sooner than its total
// wave-time over S and no sooner than its longest query, so the estimate is
// max(sum_i w * t_i / S, max_i t_i), with S the device's wave slots divided among the
// classes present. The candidate with the lowest estimate wins; a wider team must beat
// a narrower one by 3 %.
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

Trial measure_batch(ua_ctx *c, const std::vector<ua_query> &qs, int waves, double wave_slots, int *rc) {
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
    t.est_ms = std::max(sum * waves / std::max(1.0, wave_slots), mx);
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
    o->grid_large_cells = c->tune.grid_large_cells;
    o->grid_team_waves = team_waves(c, KIND_GRID, 0);
    o->grid_large_team_waves = team_waves(c, KIND_GRID, c->tune.grid_large_cells);
    o->puzzle_team_waves = team_waves(c, KIND_PUZZLE, 0);
    o->puzzle_hash_log2 = batch_hash_log2(c, 4);
    o->plain_grid_heaps = plain_heaps(c, KIND_GRID);
    o->plain_puzzle_heaps = plain_heaps(c, KIND_PUZZLE);
    o->plain_hash_log2 = plain_hash_log2(c, 4);
    return UA_OK;
}

extern "C" int ua_set_tuning(ua_ctx *c, const ua_tuning *t) {
    if (!c || !t) return fail(c, UA_ERR_INVALID, "null argument");
    auto width = [](int v) { return v >= kBlockWaves && v <= 1024 && v % kBlockWaves == 0; };
    auto in = [](int64_t v, int64_t lo, int64_t hi) { return v >= lo && v <= hi; };
    if (!width(t->grid_team_waves) || !width(t->grid_large_team_waves) || !width(t->puzzle_team_waves) ||
        !in(t->grid_large_cells, 1, 1ll << 40) || !in(t->puzzle_hash_log2, 10, 32) ||
        (t->plain_grid_heaps && !in(t->plain_grid_heaps, 4, 1 << 20)) ||
        (t->plain_puzzle_heaps && !in(t->plain_puzzle_heaps, 4, 1 << 20)) ||
        !in(t->plain_hash_log2, 10, 32))
        return fail(c, UA_ERR_INVALID, "tuning value out of range (team widths are multiples of 16 waves)");
    c->tune.grid_waves = t->grid_team_waves;
    c->tune.grid_large_waves = t->grid_large_team_waves;
    c->tune.grid_large_cells = t->grid_large_cells;
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
    std::vector<ua_query> cls[kClasses];
    for (int64_t i = 0; i < n; ++i) {
        const ua_query &u = qs[i];
        if (u.kind == UA_QUERY_GRID) {
            if (!u.grid) return fail(c, UA_ERR_INVALID, "sample query %lld: null grid", (long long)i);
            cls[class_of(c, KIND_GRID, (int64_t)u.grid->H * u.grid->W)].push_back(u);
        } else if (u.kind == UA_QUERY_PUZZLE) {
            if (u.n == 4) cls[CLS_PUZZLE].push_back(u);
        } else {
            return fail(c, UA_ERR_INVALID, "sample query %lld: bad kind", (long long)i);
        }
    }
    int per_cu = 0;
    HIPCHK(c, teams_max_blocks(&per_cu));
    const int resident = std::max(1, c->num_cu * std::max(1, per_cu));
    int present = 0;
    for (auto &v : cls) present += !v.empty();
    const double class_slots = (double)resident * kBlockWaves / std::max(1, present);
    const bool fixed[kClasses] = {c->cfg.puzzle_team_waves > 0, c->cfg.team_waves > 0, c->cfg.team_waves > 0};
    int *knob[kClasses] = {&c->tune.puzzle_waves, &c->tune.grid_waves, &c->tune.grid_large_waves};
    static const std::vector<int> cand[kClasses] = {{16, 32, 64, 128}, {16, 32}, {16, 32, 64}};
    double est[kClasses] = {0, 0, 0};
    int trials = 0, rc = 0;
    const Tuning saved = c->tune;
    auto bail = [&](int code) {
        c->tune = saved;
        return code;
    };

    // 1. team width per class; for puzzles then the hash size from the chosen width
    for (int k = 0; k < kClasses; ++k) {
        if (cls[k].empty() || fixed[k]) continue;
        int best_w = *knob[k];
        double best = 0, best_ins = 0;
        for (int w : cand[k]) {
            if (w / kBlockWaves > resident) continue;
            *knob[k] = w;
            Trial t = measure_batch(c, cls[k], w, class_slots, &rc);
            if (rc < 0) return bail(rc);
            ++trials;
            if (t.ok && (best == 0 || t.est_ms < 0.97 * best)) {
                best = t.est_ms;
                best_w = w;
                best_ins = t.max_ins;
            }
        }
        *knob[k] = best_w;
        est[k] = best;
        if (k == CLS_PUZZLE && best > 0 && c->cfg.hash_log2 <= 0) c->tune.puzzle_hash_log2[4] = hash_for(best_ins, 18, 26);
    }
    std::vector<ua_query> grids = cls[CLS_GRID];
    grids.insert(grids.end(), cls[CLS_GRID_LARGE].begin(), cls[CLS_GRID_LARGE].end());
    const std::vector<ua_query> &puzzles = cls[CLS_PUZZLE];
    // 2. heaps of the plain whole-GPU solve, per kind, on the first few sample queries
    if (c->cfg.plain_heaps <= 0) {
        const int max_heaps = kBlockWaves * resident;
        for (int kind : {KIND_GRID, KIND_PUZZLE}) {
            const std::vector<ua_query> &src = kind == KIND_GRID ? grids : puzzles;
            if (src.empty()) continue;
            const size_t m = std::min<size_t>(16, src.size());
            int best_h = 0;
            double best = 0, best_ins = 0;
            // depth-bound grids can run best on one block (fewer, fuller heaps); frontier-
            // bound puzzles want many heaps
            for (int h = kind == KIND_GRID ? kBlockWaves : 64; h <= std::min(4096, max_heaps); h *= 2) {
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
                c->tune.plain_hash_log2[4] = hash_for(best_ins, 18, 26);
        }
    }
    c->tune.tuned = true;
    rc = ua_get_tuning(c, out);
    out->trials = trials;
    out->est_grid_ms = est[CLS_GRID];
    out->est_grid_large_ms = est[CLS_GRID_LARGE];
    out->est_puzzle_ms = est[CLS_PUZZLE];
    return rc;
}
