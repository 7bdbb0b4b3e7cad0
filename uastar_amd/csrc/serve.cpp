// serve.cpp -- continuous batching of grid and puzzle queries (ua_serve_* in include/uastar.h).
//
// Submitted queries wait in a pending queue. Each ua_serve_step():
//   1. admits pending queries while a team slot is free and their memory estimate
//      (best-cost table or hash, heap storage, resume record) fits the device pool;
//   2. runs ONE bounded cooperative launch of k_teams over every resident query, in
//      admission order: each group runs queries until it has been busy for the launch's
//      time budget, then suspends the query it runs (heap tops and counters go to the
//      query's resume record; the state stays resident in HBM). A query the launch does
//      not reach waits for the next one;
//   3. retires the queries that finished (found / no path / error), frees their memory;
//   4. restarts a 15-puzzle that has run long on a wide team (a quarter of the device)
//      with a larger hash.
// Queries that arrive while others are on the device are admitted at the next step,
// so their work overlaps on the device with the queries already in flight. A heavy
// query runs for the whole launch on its team; light ones finish and make room. A query
// that outgrows its heap or hash is restarted with a larger plan.
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>

#include "internal.hpp"

using namespace ua;
using namespace uart;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device arena (offsets, 256-byte granules).
struct Pool {
    DevBuf arena;
    size_t cap = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    hipError_t init(size_t bytes) {
        bytes &= ~(size_t)255;
        hipError_t e = arena.ensure(bytes);
        if (e != hipSuccess) return e;
        cap = bytes;
        free_.clear();
        free_[0] = bytes;
        return hipSuccess;
    }
    bool alloc(size_t n, size_t *off) {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            size_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align256(n);
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

struct Job {
    uint64_t ticket = 0;
    ua_query q{};
    Clock::time_point submit;
    QueryDesc d{};
    Plan plan;
    size_t off = 0, bytes = 0;    // pool region: best table / hash, then heaps and resume record
    bool wide = false;  // planned for a wide team (CLS_WIDE) with a larger hash
    int scale = 1;      // grid heap scale / puzzle hash growth steps
    int launches = 0;
    int64_t rounds = 0;  // rounds completed so far
    double device_ms = 0, kernel_ms = 0;
};

}  // namespace

struct ServeEngine {
    std::mutex mu;  // guards pending, done, next_ticket
    std::deque<Job> pending;
    std::vector<Job> active;
    std::vector<std::pair<uint64_t, ua_result>> done;
    std::map<uint64_t, std::vector<int32_t>> paths;  // UA_FLAG_SERVE_PATHS: kept until fetched
    uint64_t next_ticket = 1;
    Pool pool;
    bool pool_ready = false;
    int64_t budget_ticks = 400000;  // device time per launch: 4 ms of s_memrealtime (100 MHz)
    double widen_after_ms = 25.0;   // a puzzle that has run this long restarts on a wide team
    int max_wide = 2;               // widened queries at a time (each takes a quarter of the device)
    int wide_hash_grow = 4;         // a widened puzzle's hash has 2^4 times the slots
    int64_t widened = 0;
    int64_t launches = 0;
    double kernel_ms = 0;
};

// the team-width class of a planned job, and of a query before planning
static int64_t cells_of(const ua_query &q) {
    return q.kind == UA_QUERY_GRID && q.grid ? (int64_t)q.grid->H * q.grid->W : 0;
}
static int cls_of_query(ua_ctx *c, const ua_query &q) {
    return class_of(c, q.kind == UA_QUERY_PUZZLE ? KIND_PUZZLE : KIND_GRID, cells_of(q));
}
static int cls_of_job(ua_ctx *c, const Job &j) { return j.wide ? CLS_WIDE : cls_of_query(c, j.q); }
// heaps of a widened query's team: a quarter of the device's resident waves
static int wide_waves(ua_ctx *c) {
    int per_cu = 0;
    if (teams_max_blocks(&per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
    return kBlockWaves * std::max(1, c->num_cu * per_cu / 4);
}

void ua_serve_release(ua_ctx *c) {
    if (!c || !c->serve) return;
    c->serve->pool.arena.release();
    delete c->serve;
    c->serve = nullptr;
}

namespace {

ServeEngine *engine(ua_ctx *c) {
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->serve) c->serve = new ServeEngine();
    return c->serve;
}

// Plans job j for its current scale and sets its device pointers inside the pool.
int plan_job(ua_ctx *c, ServeEngine *e, Job &j, bool *fits) {
    *fits = false;
    const int heaps = j.wide ? wide_waves(c)
                      : j.q.kind == UA_QUERY_PUZZLE
                          ? team_waves(c, KIND_PUZZLE, 0)
                          : team_waves(c, KIND_GRID, j.q.grid ? (int64_t)j.q.grid->H * j.q.grid->W : 0);
    int rc;
    if (j.q.kind == UA_QUERY_GRID) {
        rc = grid_fill_desc(c, j.q.grid, j.q.sx, j.q.sy, j.q.ex, j.q.ey, heaps, j.scale, j.d, j.plan);
    } else {
        rc = ensure_pdb(c, j.q.n);
        if (!rc) {
            int base = batch_hash_log2(c, j.q.n);
            int grow = j.wide ? e->wide_hash_grow : 0;
            for (int s = j.scale; s > 1; s /= 4) grow += 2;
            rc = puzzle_fill_desc(c, j.q.n, j.q.tiles, heaps, std::min(32, base + grow), j.d, j.plan);
            if (j.wide) j.d.cls = CLS_WIDE;
        }
    }
    if (rc) return rc;
    size_t rec = align256((size_t)resume_words(heaps) * 8);
    j.bytes = align256(j.plan.best_bytes) + align256(j.plan.vals_bytes) + align256(j.plan.heap_bytes) + rec;
    if (j.bytes > e->pool.cap) return fail(c, UA_ERR_OOM, "query needs %zu bytes, pool is %zu", j.bytes, e->pool.cap);
    size_t off;
    if (!e->pool.alloc(j.bytes, &off)) return UA_OK;  // wait for memory
    *fits = true;
    j.off = off;
    char *base = e->pool.arena.as<char>() + off;
    size_t o = 0;
    j.d.best = reinterpret_cast<uint64_t *>(base + o);
    o += align256(j.plan.best_bytes);
    j.d.vals = j.plan.vals_bytes ? reinterpret_cast<uint32_t *>(base + o) : nullptr;
    o += align256(j.plan.vals_bytes);
    j.d.heap = reinterpret_cast<uint64_t *>(base + o);
    o += align256(j.plan.heap_bytes);
    j.d.resume = reinterpret_cast<uint64_t *>(base + o);
    // fresh search state: grid best = all ones but blocked cells' open bit; puzzle slots = 0
    // (5x5: keys = 0, vals = all ones); no resume record
    hipError_t he;
    if (j.d.kind == KIND_GRID) {
        he = launch_grid_init_best(j.d.best, j.d.bitmap, j.d.H, j.d.W, j.d.wpr, c->st);
    } else {
        he = hipMemsetAsync(j.d.best, 0, j.plan.best_bytes, c->st);
        if (he == hipSuccess && j.plan.vals_bytes) he = hipMemsetAsync(j.d.vals, 0xFF, j.plan.vals_bytes, c->st);
    }
    if (he == hipSuccess) he = hipMemsetAsync(j.d.resume, 0, 64, c->st);
    if (he != hipSuccess) {  // the job is not admitted: give its region back
        e->pool.release(off, j.bytes);
        *fits = false;
        HIPCHK(c, he);
    }
    return UA_OK;
}

void complete(ServeEngine *e, Job &j, const ua_result &r) {
    std::lock_guard<std::mutex> g(e->mu);
    e->done.emplace_back(j.ticket, r);
}

}  // namespace

extern "C" int ua_serve_submit(ua_ctx *c, const ua_query *q, uint64_t *ticket) {
    if (!c || !q) return fail(c, UA_ERR_INVALID, "null argument");
    ServeEngine *e = engine(c);
    Job j;
    j.q = *q;
    j.submit = Clock::now();
    if (q->kind == UA_QUERY_GRID) {
        if (!q->grid) return fail(c, UA_ERR_INVALID, "grid query without a grid");
    } else if (q->kind == UA_QUERY_PUZZLE) {
        if (!board_ok(q->n, q->tiles)) return fail(c, UA_ERR_INVALID, "board is not a permutation");
    } else {
        return fail(c, UA_ERR_INVALID, "bad query kind");
    }
    std::lock_guard<std::mutex> g(e->mu);
    j.ticket = e->next_ticket++;
    if (ticket) *ticket = j.ticket;
    if (q->kind == UA_QUERY_PUZZLE && !board_solvable(q->n, q->tiles)) {
        ua_result r{};
        r.status = UA_NO_PATH;
        e->done.emplace_back(j.ticket, r);
        return UA_OK;
    }
    e->pending.push_back(j);
    return UA_OK;
}

extern "C" int ua_serve_step(ua_ctx *c, int64_t *n_active, int64_t *n_pending) {
    if (!c) return UA_ERR_INVALID;
    ServeEngine *e = engine(c);
    if (!e->pool_ready) {
        size_t free_b = 0, total_b = 0;
        HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
        size_t budget = c->cfg.mem_budget ? c->cfg.mem_budget : (size_t)(free_b * 0.80);
        HIPCHK(c, e->pool.init(budget));
        e->pool_ready = true;
    }
    // resident queries per team width: about twice the teams of that width one launch
    // could hold if it had the device to itself
    int per_cu = 0;
    HIPCHK(c, teams_max_blocks(&per_cu));
    const int resident = c->num_cu * std::max(1, per_cu);
    size_t cap_cls[kClasses], n_cls[kClasses] = {0, 0, 0};
    for (int k = 0; k < kClasses; ++k) {
        const int kind = k == CLS_PUZZLE ? KIND_PUZZLE : KIND_GRID;
        const int64_t cells = k == CLS_GRID_LARGE ? c->tune.grid_large_cells : 0;
        cap_cls[k] = (size_t)std::max(1, 2 * resident / (team_waves(c, kind, cells) / kBlockWaves));
    }
    cap_cls[CLS_WIDE] = (size_t)e->max_wide;
    for (const Job &j : e->active) n_cls[cls_of_job(c, j)]++;

    // 1. admission (in arrival order; a class at its cap waits, the other can still be admitted)
    {
        std::deque<Job> later;
        for (;;) {
            Job j;
            {
                std::lock_guard<std::mutex> g(e->mu);
                if (e->pending.empty()) break;
                j = e->pending.front();
                e->pending.pop_front();
            }
            int k = cls_of_job(c, j);
            if (n_cls[k] >= cap_cls[k]) {
                later.push_back(j);
                continue;
            }
            bool fits = false;
            int rc = plan_job(c, e, j, &fits);
            if (rc) {  // cannot ever run: report it
                ua_result r{};
                r.status = rc;
                complete(e, j, r);
                continue;
            }
            if (!fits) {  // wait for memory to free up
                later.push_back(j);
                break;
            }
            n_cls[k]++;
            e->active.push_back(j);
        }
        std::lock_guard<std::mutex> g(e->mu);
        for (auto it = later.rbegin(); it != later.rend(); ++it) e->pending.push_front(*it);
    }
    if (!e->active.empty()) {
        // 2. one bounded launch over every resident query
        std::vector<QueryDesc> descs;
        std::vector<QueryResult> init(e->active.size());
        descs.reserve(e->active.size());
        for (size_t i = 0; i < e->active.size(); ++i) {
            Job &j = e->active[i];
            j.d.round_budget = 0;
            descs.push_back(j.d);
            init[i].status = ST_SUSPENDED;  // what a query the launch does not reach reports
            init[i].rounds = j.rounds;
        }
        std::vector<QueryResult> res;
        int rc = run_batch(c, descs, res, 2, e->budget_ticks, &init);
        if (rc) return rc;
        double launch_ms = c->last_kernel_ms;
        e->launches++;
        e->kernel_ms += launch_ms;
        // 3. retirement
        std::vector<Job> still;
        auto now = Clock::now();
        for (size_t i = 0; i < e->active.size(); ++i) {
            Job &j = e->active[i];
            const QueryResult &r = res[i];
            const double dms = r.t_end > r.t_start ? (double)(r.t_end - r.t_start) / 1e5 : 0.0;
            j.launches++;
            j.kernel_ms += launch_ms;
            j.device_ms += dms;
            j.rounds = r.rounds;
            if (r.status == ST_SUSPENDED) {
                still.push_back(j);
                continue;
            }
            auto give_back = [&] {
                e->pool.release(j.off, j.bytes);
            };
            if (r.status == ST_CAPACITY && j.scale < 256) {  // restart with a larger plan
                give_back();
                j.scale *= 4;
                j.rounds = 0;
                std::lock_guard<std::mutex> g(e->mu);
                e->pending.push_front(j);
                continue;
            }
            ua_result out;
            fill_result(j.d, r, &out);
            if ((c->cfg.flags & UA_FLAG_SERVE_PATHS) && out.status == UA_OK) {
                // the walk reads the query's table, which goes back to the pool below
                std::vector<int32_t> path;
                const int prc = fetch_path(c, j.d, r, path);
                if (prc) {  // the query fails with the walk's error (ua_last_error has it)
                    out.status = prc;
                    out.found = 0;
                } else {
                    std::lock_guard<std::mutex> g(e->mu);
                    e->paths[j.ticket] = std::move(path);
                }
            }
            give_back();
            out.kernel_ms = j.kernel_ms;
            out.device_ms = j.device_ms;
            out.latency_ms = std::chrono::duration<double, std::milli>(now - j.submit).count();
            complete(e, j, out);
        }
        e->active.swap(still);
        // 4. widen long-running puzzles (frontier-bound: more heaps pop more per round). A
        // puzzle that has run widen_after_ms restarts from its start board on a team of a
        // quarter of the device with a hash of 2^wide_hash_grow times the slots: a long
        // search on a wide team generates states several times faster, and a hash near
        // full lengthens every probe and ends in a capacity restart anyway. The narrow
        // team's work is lost, about an eighth of what the wide team does in the same time.
        const int ww = wide_waves(c);
        int n_wide = 0;
        for (const Job &j : e->active) n_wide += j.wide;
        {
            std::lock_guard<std::mutex> g(e->mu);
            for (const Job &j : e->pending) n_wide += j.wide;
        }
        std::vector<Job> keep;
        std::vector<Job> restart;
        for (Job &j : e->active) {
            if (n_wide < e->max_wide && !j.wide && j.q.kind == UA_QUERY_PUZZLE && j.q.n == 4 &&
                j.device_ms >= e->widen_after_ms && j.d.heaps < ww) {
                e->pool.release(j.off, j.bytes);
                j.wide = true;
                j.rounds = 0;
                ++n_wide;
                ++e->widened;
                restart.push_back(j);
            } else {
                keep.push_back(j);
            }
        }
        e->active.swap(keep);
        if (!restart.empty()) {
            std::lock_guard<std::mutex> g(e->mu);
            for (auto it = restart.rbegin(); it != restart.rend(); ++it) e->pending.push_front(*it);
        }
    }
    if (n_active) *n_active = (int64_t)e->active.size();
    if (n_pending) {
        std::lock_guard<std::mutex> g(e->mu);
        *n_pending = (int64_t)e->pending.size();
    }
    return UA_OK;
}

extern "C" int ua_serve_poll(ua_ctx *c, uint64_t *tickets, ua_result *results, int64_t cap, int64_t *n_out) {
    if (!c || !n_out || cap < 0) return fail(c, UA_ERR_INVALID, "bad arguments");
    ServeEngine *e = engine(c);
    std::lock_guard<std::mutex> g(e->mu);
    int64_t n = std::min<int64_t>(cap, (int64_t)e->done.size());
    for (int64_t i = 0; i < n; ++i) {
        if (tickets) tickets[i] = e->done[i].first;
        if (results) results[i] = e->done[i].second;
    }
    e->done.erase(e->done.begin(), e->done.begin() + n);
    *n_out = n;
    return UA_OK;
}

extern "C" int ua_serve_path(ua_ctx *c, uint64_t ticket, int32_t *out, int64_t cap, int64_t *len) {
    if (!c || !len) return fail(c, UA_ERR_INVALID, "null argument");
    if (!(c->cfg.flags & UA_FLAG_SERVE_PATHS))
        return fail(c, UA_ERR_INVALID, "served paths are kept only with UA_FLAG_SERVE_PATHS");
    ServeEngine *e = engine(c);
    std::lock_guard<std::mutex> g(e->mu);
    auto it = e->paths.find(ticket);
    if (it == e->paths.end()) return fail(c, UA_ERR_INVALID, "no path kept for ticket %llu", (unsigned long long)ticket);
    *len = (int64_t)it->second.size();
    if (!out) return UA_OK;
    if (*len > cap) return fail(c, UA_ERR_INVALID, "path needs %lld entries", (long long)*len);
    std::copy(it->second.begin(), it->second.end(), out);
    e->paths.erase(it);
    return UA_OK;
}

extern "C" int ua_serve_drain(ua_ctx *c) {
    if (!c) return UA_ERR_INVALID;
    for (;;) {
        int64_t a = 0, p = 0;
        int rc = ua_serve_step(c, &a, &p);
        if (rc) return rc;
        if (!a && !p) return UA_OK;
    }
}

extern "C" int ua_serve_stats(ua_ctx *c, int64_t *launches, double *kernel_ms, int32_t *budget_us,
                              int64_t *widened) {
    if (!c) return UA_ERR_INVALID;
    ServeEngine *e = engine(c);
    if (launches) *launches = e->launches;
    if (kernel_ms) *kernel_ms = e->kernel_ms;
    if (budget_us) *budget_us = (int32_t)(e->budget_ticks / 100);
    if (widened) *widened = e->widened;
    return UA_OK;
}
