// runtime.cpp -- host side of libuastar: the C ABI of include/uastar.h.
//
// A ua_ctx owns one HIP stream and every device allocation. The plain solve
// (ua_grid_solve / ua_puzzle_solve) runs one query on a cooperative launch whose
// waves all work on it; batches and served queries run one query per workgroup.
// Device state is never reset (uAstar calls cudaDeviceReset in every initialize,
// pathway/GPU-solver.cu:93-94); buffers grow and are reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace ua;
using namespace uart;

// Scratch of ua_grid_solve_noalloc: one array per GPU, held until the process ends. Host
// code uses it only through the device address hipGetSymbolAddress gives at ua_create
// (ua_ctx::noalloc_scratch); the name itself is a host-side stand-in. Layout, for a grid of
// at most kNoAllocCells cells: the launch's I/O block, the path walk's job and ids, the best
// table (8 B a cell), then the bucket ring (its counters and 256 x 2^kNoAllocBktLog2 keys).
namespace {
constexpr int64_t kNoAllocCells = 1ll << 20;
constexpr int kNoAllocBktLog2 = 14;
constexpr size_t kNoAllocIo = 16384;                                    // I/O block and path job
constexpr size_t kNoAllocPath = (size_t)kNoAllocCells * 4 + 256;        // path ids
constexpr size_t kNoAllocBest = (size_t)kNoAllocCells * 8;              // best table
constexpr size_t kNoAllocRing = ((size_t)kBucketMetaWords + ((size_t)kBuckets << kNoAllocBktLog2)) * 8;
constexpr size_t kNoAllocBytes = kNoAllocIo + kNoAllocPath + kNoAllocBest + kNoAllocRing;
}  // namespace
static __device__ __align__(256) uint64_t scratch[(kNoAllocBytes + 7) / 8];

namespace uart {

int fail(ua_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

int status_of(const QueryResult &r) {
    if (r.status == ST_OK) return r.found ? UA_OK : UA_NO_PATH;
    return r.status;
}

// ---- per-query planning -------------------------------------------------------

int grid_fill_desc(ua_ctx *c, const ua_grid *g, int sx, int sy, int ex, int ey, int heaps, int heap_scale,
                   QueryDesc &q, Plan &p) {
    if (!g || !g->bm) return fail(c, UA_ERR_INVALID, "null grid");
    if (sx < 0 || sx >= g->H || sy < 0 || sy >= g->W || ex < 0 || ex >= g->H || ey < 0 || ey >= g->W)
        return fail(c, UA_ERR_INVALID, "start/goal out of range");
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_GRID;
    q.H = g->H;
    q.W = g->W;
    q.sx = sx;
    q.sy = sy;
    q.ex = ex;
    q.ey = ey;
    q.wpr = g->wpr;
    q.ybits = bits_for(g->W);
    q.cbits = bits_for(g->H) + q.ybits;
    q.qshift = std::max(0, q.cbits - 10);
    q.bitmap = g->bm;
    q.heaps = heaps;
    q.cls = class_of(c, KIND_GRID, (int64_t)g->H * g->W);
    q.max_rounds = c->cfg.max_rounds > 0 ? c->cfg.max_rounds : (1 << 26);
    int64_t cells = (int64_t)g->H * g->W;
    int64_t items = std::max<int64_t>(65536, cells / 8) * heap_scale;  // open-list capacity
    p = Plan{};
    p.best_bytes = (size_t)cells * 8;
    if (heaps == kBlockWaves && !(c->cfg.flags & UA_FLAG_GRID_HEAPS)) {
        // a one-block team: the bucket ring (a ring shared across blocks measured slower
        // than their heaps: agent-scope counters and two grid barriers a round). Buckets
        // are 1/16 step of f wide (bshift), a round takes keys from the lowest bwin of them
        // (at most kBuckets - 47: a child's f exceeds its parent's by at most 2 sqrt(2)
        // steps, 46 buckets); each holds 1/8 of the open-list capacity, at least 4096 keys.
        // A grid whose optimum is its octile distance puts every cell of every optimal
        // path in one bucket (f = C* exactly): the serving stream's deepest grid, 7786^2,
        // overflowed 1/32. A bucket that fills up fails the query with ST_CAPACITY, which
        // retries it larger
        q.open_kind = OPEN_BUCKETS;
        q.bshift = std::max(0, 32 - q.qshift);
        q.bwin = std::max(1, std::min(c->tune.grid_bucket_window, kBuckets - 48));
        q.bkt_log2 = bits_for(std::max<int64_t>(4096, items / 8));
        q.heap_cap = 0;
        p.heap_bytes = ((size_t)kBucketMetaWords + ((size_t)kBuckets << q.bkt_log2)) * 8;  // counters, then keys
        return UA_OK;
    }
    int64_t per_heap = (items + 64ll * heaps - 1) / (64ll * heaps);
    q.heap_cap = (int)std::min<int64_t>(std::max<int64_t>(per_heap + 1, 4), 1 << 30);
    p.heap_bytes = (size_t)heaps * (size_t)(q.heap_cap - 1) * 512;
    return UA_OK;
}

bool board_ok(int n, const uint8_t *t) {
    if (n < 3 || n > 5) return false;
    std::vector<int> seen(n * n, 0);
    for (int i = 0; i < n * n; ++i) {
        if (t[i] >= n * n || seen[t[i]]) return false;
        seen[t[i]] = 1;
    }
    return true;
}

// parity of the permutation of the n*n squares (blank = n*n) equals the parity
// of the blank's Manhattan distance to the bottom-right cell
bool board_solvable(int n, const uint8_t *t) {
    int nn = n * n;
    std::vector<int> p(nn);
    int blank = 0;
    for (int i = 0; i < nn; ++i) {
        p[i] = t[i] ? t[i] - 1 : nn - 1;
        if (!t[i]) blank = i;
    }
    std::vector<char> seen(nn, 0);
    int cycles = 0;
    for (int i = 0; i < nn; ++i) {
        if (seen[i]) continue;
        ++cycles;
        for (int j = i; !seen[j]; j = p[j]) seen[j] = 1;
    }
    int man = (n - 1 - blank / n) + (n - 1 - blank % n);
    return ((nn - cycles) & 1) == (man & 1);
}

uint64_t pack_board(int n, const uint8_t *t) {
    uint64_t s = 0;
    for (int i = 0; i < n * n; ++i) s |= (uint64_t)t[i] << (4 * i);
    return s;
}

// 5x5 boards as PuzzleStorage<5> (puzzle/storage.hpp:144-170): cells 0..11 in w0, cells
// 12..23 in w1, five bits each; cell 24's low four bits at w0 bit 60, its bit 4 at w1 bit 60
void pack_board5(const uint8_t *t, uint64_t *w0, uint64_t *w1) {
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 12; ++i) a |= (uint64_t)t[i] << (5 * i);
    for (int i = 12; i < 24; ++i) b |= (uint64_t)t[i] << (5 * (i - 12));
    a |= (uint64_t)(t[24] & 15) << 60;
    b |= (uint64_t)(t[24] & 16) << 56;
    *w0 = a;
    *w1 = b;
}

int puzzle_fill_desc(ua_ctx *c, int n, const uint8_t *tiles, int heaps, int hash_log2, QueryDesc &q, Plan &p) {
    if (hash_log2 < 1 || hash_log2 > kPuzzleSlotBits)
        return fail(c, UA_ERR_INVALID, "puzzle hash of 2^%d slots: the item key holds %d slot bits", hash_log2,
                    kPuzzleSlotBits);
    std::memset(&q, 0, sizeof q);
    q.kind = KIND_PUZZLE;
    q.n = n;
    if (n == 5) pack_board5(tiles, &q.start_state, &q.start_state2);
    else q.start_state = pack_board(n, tiles);
    q.pdb = c->pdb[n].as<uint8_t>();
    q.hash_log2 = hash_log2;
    q.heaps = heaps;
    q.cls = CLS_PUZZLE;
    q.max_rounds = c->cfg.max_rounds > 0 ? c->cfg.max_rounds : (1 << 26);
    int64_t items = std::max<int64_t>(65536, (1ll << hash_log2) / 2);
    int64_t per_heap = (items + 64ll * heaps - 1) / (64ll * heaps);
    q.heap_cap = (int)std::max<int64_t>(per_heap + 1, 4);
    p = Plan{};
    // n = 5: two key words a slot and a value array; n = 3, 4: 16-byte slots of board and
    // complemented value, empty when zeroed
    p.best_bytes = (size_t)16 << hash_log2;
    p.vals_bytes = n == 5 ? (size_t)4 << hash_log2 : 0;
    p.heap_bytes = (size_t)heaps * (size_t)(q.heap_cap - 1) * 512;
    return UA_OK;
}

// the partitions' tables of side n, concatenated: 9!, 16!/8! + 16!/9!, 4 x 25!/19!
static uint64_t pdb_bytes(int n) { return n == 5 ? 510048000ull : n == 4 ? 576576000ull : 362880ull; }

int ensure_pdb(ua_ctx *c, int n) {
    if (n < 3 || n > 5) return fail(c, UA_ERR_INVALID, "puzzle side must be 3, 4 or 5");
    if (c->pdb[n].p) return UA_OK;
    HIPCHK(c, c->pdb[n].ensure(pdb_bytes(n)));
    HIPCHK(c, c->d_added.ensure(sizeof(unsigned long long)));
    hipError_t e = build_pdb(n, c->pdb[n].as<uint8_t>(), c->d_added.as<unsigned long long>(), c->st);
    if (e != hipSuccess) {
        c->pdb[n].release();
        return fail(c, UA_ERR_HIP, "PDB build: %s", hipGetErrorString(e));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return UA_OK;
}

void fill_result(const QueryDesc &q, const QueryResult &r, ua_result *o) {
    std::memset(o, 0, sizeof *o);
    o->status = status_of(r);
    o->found = r.found && r.status == ST_OK;
    if (o->found) {
        if (q.kind == KIND_GRID) {
            int64_t a, b;
            key_to_pair(r.goal, &a, &b);
            o->orth = a;
            o->diag = b;
            o->cost = (double)a + (double)b * SQRT2;
        } else {
            o->steps = (int64_t)(r.goal >> 32);
            o->cost = (double)o->steps;
        }
    }
    o->heaps = q.heaps;
    o->rounds = r.rounds;
    o->pops = r.pops;
    o->expanded = r.expanded;
    o->generated = r.generated;
    o->inserted = r.inserted;
    o->atomics = r.atomics;
    o->device_ms = ticks_ms(r.t_start, r.t_end);
    o->weight = 1 << q.wshift;
    o->optimal = o->found && q.wshift == 0 && !(r.flags & FLAG_PRUNED);
}

// ---- plain solve (whole GPU, cooperative launch) ----------------------------------

// Heaps of the plain whole-GPU solve for a query kind (KIND_GRID / KIND_PUZZLE): the
// configured or tuned count, rounded to whole 4-wave blocks and capped at what one
// cooperative launch can hold.
int plain_heaps(ua_ctx *c, int kind, int64_t cells) {
    int per_cu = 0;
    if (teams_max_blocks(&per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
    int tuned = c->tune.plain_heaps[kind == KIND_PUZZLE ? 1 : 0];
    // untuned defaults: a grid below 2^22 cells gets one block (16 waves) on the bucket
    // ring, whose rounds are cheapest (1000^2: 10 ms against 16 ms on 256 heaps); a larger
    // grid 256 heaps (16 blocks), whose throughput a one-block team cannot match (16384^2:
    // 195 M expansions); a puzzle 4 heaps per CU, whose open list is wide
    const int dflt = kind == KIND_PUZZLE ? 4 * c->num_cu
                                         : cells < (1ll << 22) ? kBlockWaves : std::min(256, 4 * c->num_cu);
    int want = c->cfg.plain_heaps > 0 ? c->cfg.plain_heaps : (tuned > 0 ? tuned : dflt);
    int blocks = std::max(1, (want + kBlockWaves - 1) / kBlockWaves);
    blocks = std::min(blocks, per_cu * c->num_cu);
    return blocks * kBlockWaves;
}

// Hash slots (log2) per puzzle query: batched / served, and plain.
int batch_hash_log2(ua_ctx *c, int n) {
    return c->cfg.hash_log2 > 0 ? c->cfg.hash_log2 : c->tune.puzzle_hash_log2[n >= 3 && n <= 5 ? n : 4];
}
int plain_hash_log2(ua_ctx *c, int n) {
    return c->cfg.hash_log2 > 0 ? c->cfg.hash_log2 : c->tune.plain_hash_log2[n >= 3 && n <= 5 ? n : 4];
}

static int run_wide(ua_ctx *c, QueryDesc &q, const Plan &p) {
    int heaps = q.heaps;
    int blocks = heaps / kBlockWaves;
    int segcap = heaps * 64 * (q.kind == KIND_GRID ? 8 : 4);
    HIPCHK(c, c->best.ensure(p.best_bytes));
    if (p.vals_bytes) HIPCHK(c, c->vals.ensure(p.vals_bytes));
    HIPCHK(c, c->heap.ensure(p.heap_bytes));
    q.best = c->best.as<uint64_t>();
    q.vals = c->vals.as<uint32_t>();
    q.heap = c->heap.as<uint64_t>();
    if (q.kind == KIND_GRID) {
        HIPCHK(c, launch_grid_init_best(q.best, q.bitmap, q.H, q.W, q.wpr, c->st));
    } else {
        HIPCHK(c, hipMemsetAsync(q.best, 0, p.best_bytes, c->st));
        if (p.vals_bytes) HIPCHK(c, hipMemsetAsync(q.vals, 0xFF, p.vals_bytes, c->st));
    }
    LaunchShape shape{};
    shape.groups[0] = 1;
    shape.bpg[0] = blocks;
    shape.nq[0] = 1;
    shape.segcap[0] = segcap;
    for (int k = 1; k < kClasses; ++k) shape.bpg[k] = 1;
    QueryResult r{};
    int rc = run_step(c, 0, set_of(q), blocks, shape, &q, nullptr, 1, &r);
    if (rc) return rc;
    c->last_q = q;
    c->last_r = r;
    c->last_valid = true;
    return UA_OK;
}

}  // namespace uart

// =================================================================== C ABI

extern "C" int ua_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int ua_create(int ordinal, const ua_config *cfg, ua_ctx **out) {
    if (!out) return UA_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return UA_ERR_NO_DEVICE;
    if (ordinal < 0 || ordinal >= n) return UA_ERR_INVALID;
    ua_ctx *c = new ua_ctx();
    c->dev = ordinal;
    if (cfg) c->cfg = *cfg;
    c->graphs = (c->cfg.flags & UA_FLAG_GRAPH) != 0;
    bool ok = hipSetDevice(ordinal) == hipSuccess &&
              hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
    if (!ok) {
        ua_destroy(c);
        return UA_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) == hipSuccess) {
        c->num_cu = prop.multiProcessorCount;
        c->total_mem = prop.totalGlobalMem;
    }
    if (c->num_cu <= 0) c->num_cu = 1;
    // this device's copy of the allocation-free solve's scratch array (each GPU holds its
    // own, at its own address); without it that entry point reports an error
    void *sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess) c->noalloc_scratch = static_cast<char *>(sp);
    else (void)hipGetLastError();
    *out = c;
    return UA_OK;
}

extern "C" void ua_destroy(ua_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->dev);
    (void)hipStreamSynchronize(c->st);
    ua_serve_release(c);
    for (auto &b : c->pdb) b.release();
    release_steps(c);
    for (DevBuf *b : {&c->d_added, &c->best, &c->vals, &c->heap, &c->pathbuf}) b->release();
    for (auto &b : c->b_arenas) b.release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

extern "C" const char *ua_last_error(ua_ctx *c) { return c ? c->err.c_str() : "no context"; }

extern "C" int ua_grid_create(ua_ctx *c, int H, int W, const uint64_t *host_bitmap, ua_grid **out) {
    if (!c || !out || H <= 0 || W <= 0) return fail(c, UA_ERR_INVALID, "bad grid size");
    if ((int64_t)H * W > (1ll << 31)) return fail(c, UA_ERR_INVALID, "grid larger than 2^31 cells");
    ua_grid *g = new ua_grid{H, W, (W + 63) / 64, nullptr};
    size_t bytes = (size_t)H * g->wpr * 8;
    if (hipMalloc(&g->bm, bytes) != hipSuccess) {
        delete g;
        return fail(c, UA_ERR_OOM, "grid bitmap allocation (%zu bytes)", bytes);
    }
    if (host_bitmap) {
        hipError_t e = hipMemcpyAsync(g->bm, host_bitmap, bytes, hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) {
            (void)hipFree(g->bm);
            delete g;
            return fail(c, UA_ERR_HIP, "bitmap upload: %s", hipGetErrorString(e));
        }
    }
    *out = g;
    return UA_OK;
}

extern "C" int ua_grid_from_bytes(ua_ctx *c, int H, int W, const uint8_t *bytes, ua_grid **out) {
    if (!bytes) return fail(c, UA_ERR_INVALID, "null bytes");
    int rc = ua_grid_create(c, H, W, nullptr, out);
    if (rc) return rc;
    const size_t n = (size_t)H * W, flag_off = (n + 255) & ~(size_t)255;
    DevBuf tmp;  // the bytes, then the partial-mask flag
    uint32_t bad = 0;
    hipError_t e = tmp.ensure(flag_off + 4);
    if (!e) e = hipMemcpyAsync(tmp.p, bytes, n, hipMemcpyHostToDevice, c->st);
    if (!e) e = hipMemsetAsync(tmp.as<char>() + flag_off, 0, 4, c->st);
    if (!e) e = launch_grid_pack(tmp.as<uint8_t>(), (*out)->bm, H, W, (*out)->wpr,
                                 reinterpret_cast<uint32_t *>(tmp.as<char>() + flag_off), c->st);
    if (!e) e = hipMemcpyAsync(&bad, tmp.as<char>() + flag_off, 4, hipMemcpyDeviceToHost, c->st);
    if (!e) e = hipStreamSynchronize(c->st);
    tmp.release();
    if (e || bad) {
        ua_grid_free(c, *out);
        *out = nullptr;
        if (e) return fail(c, e == hipErrorOutOfMemory ? UA_ERR_OOM : UA_ERR_HIP, "grid upload: %s",
                           hipGetErrorString(e));
        return fail(c, UA_ERR_INVALID,
                    "cell bytes must be 0 (blocked) or 0xFF (open): a partial out-move mask cannot be held "
                    "by the 1-bit grid");
    }
    return UA_OK;
}

extern "C" int ua_grid_generate(ua_ctx *c, int H, int W, uint64_t seed, double block_rate, int sx, int sy,
                                int ex, int ey, ua_grid **out) {
    if (!(block_rate >= 0.0 && block_rate <= 1.0)) return fail(c, UA_ERR_INVALID, "block_rate outside [0,1]");
    int rc = ua_grid_create(c, H, W, nullptr, out);
    if (rc) return rc;
    uint64_t thr = (uint64_t)std::floor(block_rate * 4294967296.0);
    HIPCHK(c, launch_grid_gen((*out)->bm, H, W, (*out)->wpr, seed, thr, sx, sy, ex, ey, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return UA_OK;
}

extern "C" int ua_grid_download(ua_ctx *c, const ua_grid *g, uint64_t *host) {
    if (!c || !g || !host) return fail(c, UA_ERR_INVALID, "null argument");
    HIPCHK(c, hipMemcpyAsync(host, g->bm, (size_t)g->H * g->wpr * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return UA_OK;
}

extern "C" void ua_grid_free(ua_ctx *c, ua_grid *g) {
    if (!g) return;
    if (c) (void)hipStreamSynchronize(c->st);
    if (g->bm) (void)hipFree(g->bm);
    delete g;
}

extern "C" int ua_grid_solve(ua_ctx *c, const ua_grid *g, int sx, int sy, int ex, int ey, ua_result *res) {
    if (!c || !res) return fail(c, UA_ERR_INVALID, "null argument");
    c->last_valid = false;
    if (!g) return fail(c, UA_ERR_INVALID, "null grid");
    int heaps = plain_heaps(c, KIND_GRID, (int64_t)g->H * g->W);
    for (int scale = 1; scale <= 64; scale *= 4) {
        QueryDesc q;
        Plan p;
        int rc = grid_fill_desc(c, g, sx, sy, ex, ey, heaps, scale, q, p);
        if (rc) return rc;
        rc = run_wide(c, q, p);
        if (rc) return rc;
        if (c->last_r.status != ST_CAPACITY) break;
    }
    fill_result(c->last_q, c->last_r, res);
    res->kernel_ms = c->last_kernel_ms;
    if (res->status < 0) fail(c, res->status, "grid solve failed with status %d", res->status);
    return res->status;
}

namespace uart {

// Paths of finished searches whose state is still on the device: start..goal cell ids
// (grid) or blank moves (puzzle), walked from each goal's parent directions in one
// launch, one lane per query. A found answer fixes the length (orth + diag steps, or the
// move count), so each walk's output is exactly that long. Queries without a path get an
// empty vector.
int fetch_paths(ua_ctx *c, int n, const QueryDesc *const *qs, const QueryResult *const *rs,
                std::vector<int32_t> *out) {
    std::vector<PathJob> jobs;
    std::vector<int> which;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        out[i].clear();
        const QueryDesc &q = *qs[i];
        const QueryResult &r = *rs[i];
        if (!r.found || r.status != ST_OK) continue;
        int64_t want;
        if (q.kind == KIND_GRID) {
            int64_t a, b;
            key_to_pair(r.goal, &a, &b);
            want = a + b + 1;
            if (want > (int64_t)q.H * q.W) return fail(c, UA_ERR_HIP, "implausible path length %lld", (long long)want);
        } else {
            if (r.rounds == 0) continue;  // start == goal: no moves
            want = (int64_t)(r.goal >> 32);
            if (want > 4096) return fail(c, UA_ERR_HIP, "implausible move count %lld", (long long)want);
        }
        PathJob j{};
        j.best = q.best;
        j.vals = q.vals;
        j.kind = q.kind;
        j.n = q.n;
        j.H = q.H;
        j.W = q.W;
        j.sx = q.sx;
        j.sy = q.sy;
        j.ex = q.ex;
        j.ey = q.ey;
        j.hash_log2 = q.hash_log2;
        j.goal_slot = r.goal & 0xffffffffull;
        j.off = total;
        j.len = want;
        total += want;
        jobs.push_back(j);
        which.push_back(i);
    }
    if (jobs.empty()) return UA_OK;
    const size_t jb = align256(jobs.size() * sizeof(PathJob));
    HIPCHK(c, c->pathbuf.ensure(jb + (size_t)total * 4));
    PathJob *d_jobs = c->pathbuf.as<PathJob>();
    int32_t *d_ids = reinterpret_cast<int32_t *>(c->pathbuf.as<char>() + jb);
    HIPCHK(c, hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(PathJob), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_paths(d_jobs, (int)jobs.size(), d_ids, c->st));
    std::vector<int32_t> ids((size_t)total);
    HIPCHK(c, hipMemcpyAsync(jobs.data(), d_jobs, jobs.size() * sizeof(PathJob), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ids.data(), d_ids, (size_t)total * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t k = 0; k < jobs.size(); ++k) {
        const PathJob &j = jobs[k];
        if (j.got != j.len)
            return fail(c, UA_ERR_HIP, "broken parent chain (%lld of %lld)", (long long)j.got, (long long)j.len);
        out[which[k]].assign(ids.rbegin() + (total - j.off - j.len), ids.rbegin() + (total - j.off));
    }
    return UA_OK;
}

int fetch_path(ua_ctx *c, const QueryDesc &q, const QueryResult &r, std::vector<int32_t> &out) {
    if (!r.found || r.status != ST_OK) return fail(c, UA_ERR_INVALID, "query has no path");
    const QueryDesc *qp = &q;
    const QueryResult *rp = &r;
    return fetch_paths(c, 1, &qp, &rp, &out);
}

}  // namespace uart

// A path into a caller buffer: *len gets its length; with ids == nullptr only the length.
static int copy_path(ua_ctx *c, const std::vector<int32_t> &p, int32_t *ids, int64_t cap, int64_t *len) {
    *len = (int64_t)p.size();
    if (!ids) return UA_OK;
    if ((int64_t)p.size() > cap) return fail(c, UA_ERR_INVALID, "path needs %lld entries", (long long)p.size());
    std::copy(p.begin(), p.end(), ids);
    return UA_OK;
}

extern "C" int ua_grid_path(ua_ctx *c, int32_t *ids, int64_t cap, int64_t *len) {
    if (!c || !len) return fail(c, UA_ERR_INVALID, "null argument");
    if (!c->last_valid || c->last_q.kind != KIND_GRID || !c->last_r.found)
        return fail(c, UA_ERR_INVALID, "no solved grid query");
    std::vector<int32_t> p;
    int rc = fetch_path(c, c->last_q, c->last_r, p);
    if (rc) return rc;
    return copy_path(c, p, ids, cap, len);
}

// The plain grid solve and its path with no allocation: every device buffer of the call is
// a region of the scratch array, and the I/O block is staged from the host stack. One
// block (16 waves) on the bucket ring, as ua_grid_solve runs grids below 2^22 cells untuned.
// Calls that overlap on one GPU share the array and race.
extern "C" int ua_grid_solve_noalloc(ua_ctx *c, const ua_grid *g, int sx, int sy, int ex, int ey, ua_result *res,
                                     int32_t *path, int64_t cap, int64_t *len) {
    if (!c || !g || !res) return fail(c, UA_ERR_INVALID, "null argument");
    c->last_valid = false;
    std::memset(res, 0, sizeof *res);
    if (len) *len = 0;
    const int64_t cells = (int64_t)g->H * g->W;
    if (cells > kNoAllocCells)
        return fail(c, UA_ERR_INVALID, "ua_grid_solve_noalloc holds grids of at most %lld cells (this one has %lld)",
                    (long long)kNoAllocCells, (long long)cells);
    if (!c->noalloc_scratch) return fail(c, UA_ERR_HIP, "the scratch array's device address is unknown");
    QueryDesc q;
    Plan p;
    int rc = grid_fill_desc(c, g, sx, sy, ex, ey, kBlockWaves, 1, q, p);
    if (rc) return rc;
    char *const base = c->noalloc_scratch;
    char *const io = base;
    int32_t *const d_path = reinterpret_cast<int32_t *>(base + kNoAllocIo);
    q.best = reinterpret_cast<uint64_t *>(base + kNoAllocIo + kNoAllocPath);
    q.heap = reinterpret_cast<uint64_t *>(base + kNoAllocIo + kNoAllocPath + kNoAllocBest);
    q.vals = nullptr;
    q.open_kind = OPEN_BUCKETS;  // whatever the context's flags: the scratch holds a ring
    q.bshift = std::max(0, 32 - q.qshift);
    q.bwin = std::max(1, std::min(c->tune.grid_bucket_window, kBuckets - 48));
    q.bkt_log2 = kNoAllocBktLog2;
    q.heap_cap = 0;
    HIPCHK(c, launch_grid_init_best(q.best, q.bitmap, q.H, q.W, q.wpr, c->st));
    // the I/O block of one one-block group: header, TeamVars (zero), descriptor, result
    const size_t tv_off = align256(sizeof(StepHeader));
    const size_t qd_off = tv_off + align256(sizeof(TeamVars));
    const size_t qr_off = qd_off + align256(sizeof(QueryDesc));
    const size_t end = qr_off + sizeof(QueryResult);
    static_assert(3 * 256 + sizeof(StepHeader) + sizeof(TeamVars) + sizeof(QueryDesc) + sizeof(QueryResult) +
                          sizeof(PathJob) + 256 <= kNoAllocIo,
                  "the I/O block and the path job fit their region");
    alignas(256) char blk[kNoAllocIo];
    std::memset(blk, 0, end);
    StepHeader h{};
    h.shape.groups[0] = 1;
    for (int k = 0; k < kClasses; ++k) h.shape.bpg[k] = 1;
    h.shape.nq[0] = 1;
    h.shape.segcap[0] = kBlockWaves * 64 * 8;  // the LDS buffer's size: the group keeps it in LDS
    h.tv_off = (uint32_t)tv_off;
    h.qd_off = (uint32_t)qd_off;
    h.qr_off = (uint32_t)qr_off;
    h.nq = 1;
    std::memcpy(blk, &h, sizeof h);
    std::memcpy(blk + qd_off, &q, sizeof q);
    HIPCHK(c, hipMemcpyAsync(io, blk, end, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipEventRecord(c->ev0, c->st));
    HIPCHK(c, launch_teams(io, 1, reinterpret_cast<uint64_t *>(base), 0, SET_MAIN, c->st));
    HIPCHK(c, hipEventRecord(c->ev1, c->st));
    QueryResult r{};
    HIPCHK(c, hipMemcpyAsync(&r, io + qr_off, sizeof r, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float kms = 0;
    HIPCHK(c, hipEventElapsedTime(&kms, c->ev0, c->ev1));
    fill_result(q, r, res);
    res->kernel_ms = kms;
    if (res->status < 0) return fail(c, res->status, "grid solve failed with status %d", res->status);
    if (res->status != UA_OK || !len) return res->status;
    // the path: one walk into the scratch, then reversed into the caller's buffer
    const int64_t want = res->orth + res->diag + 1;
    if (want > cap && path) return fail(c, UA_ERR_INVALID, "path needs %lld entries", (long long)want);
    *len = want;
    if (!path) return res->status;
    PathJob j{};
    j.best = q.best;
    j.kind = KIND_GRID;
    j.H = q.H;
    j.W = q.W;
    j.sx = q.sx;
    j.sy = q.sy;
    j.ex = q.ex;
    j.ey = q.ey;
    j.len = want;
    PathJob *const d_job = reinterpret_cast<PathJob *>(io + align256(end));
    HIPCHK(c, hipMemcpyAsync(d_job, &j, sizeof j, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_paths(d_job, 1, d_path, c->st));
    HIPCHK(c, hipMemcpyAsync(&j, d_job, sizeof j, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(path, d_path, (size_t)want * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (j.got != want) return fail(c, UA_ERR_HIP, "broken parent chain (%lld of %lld)", (long long)j.got, (long long)want);
    std::reverse(path, path + want);  // the walk ran goal -> start
    return res->status;
}

extern "C" int ua_puzzle_build_pdb(ua_ctx *c, int n) {
    if (!c) return UA_ERR_INVALID;
    return ensure_pdb(c, n);
}

extern "C" int ua_puzzle_load_pdb(ua_ctx *c, int n, const uint8_t *host, uint64_t size) {
    if (!c || !host) return fail(c, UA_ERR_INVALID, "null argument");
    if (n < 3 || n > 5) return fail(c, UA_ERR_INVALID, "puzzle side must be 3, 4 or 5");
    if (size != pdb_bytes(n)) return fail(c, UA_ERR_INVALID, "a %dx%d PDB holds %llu bytes", n, n,
                                          (unsigned long long)pdb_bytes(n));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->pdb[n].ensure(pdb_bytes(n)));
    HIPCHK(c, hipMemcpy(c->pdb[n].p, host, pdb_bytes(n), hipMemcpyHostToDevice));
    return UA_OK;
}

extern "C" int ua_puzzle_download_pdb(ua_ctx *c, int n, uint8_t *host, uint64_t size) {
    if (!c || !host) return fail(c, UA_ERR_INVALID, "null argument");
    int rc = ensure_pdb(c, n);
    if (rc) return rc;
    if (size < pdb_bytes(n)) return fail(c, UA_ERR_INVALID, "buffer too small");
    HIPCHK(c, hipMemcpy(host, c->pdb[n].p, pdb_bytes(n), hipMemcpyDeviceToHost));
    return UA_OK;
}

extern "C" int ua_puzzle_solve(ua_ctx *c, int n, const uint8_t *tiles, ua_result *res) {
    if (!c || !res || !tiles) return fail(c, UA_ERR_INVALID, "null argument");
    c->last_valid = false;
    std::memset(res, 0, sizeof *res);
    if (!board_ok(n, tiles)) return fail(c, UA_ERR_INVALID, "board is not a permutation of 0..n*n-1");
    if (!board_solvable(n, tiles)) {
        res->status = UA_NO_PATH;
        return UA_NO_PATH;
    }
    int rc = ensure_pdb(c, n);
    if (rc) return rc;
    int heaps = plain_heaps(c, KIND_PUZZLE);
    int hl = plain_hash_log2(c, n);
    // a puzzle that fills its hash is a heavy one: each retry (a rerun from the start) gets
    // 16x the slots, which spares the second retry that 4x often needed
    double kms = 0;  // every attempt's launch
    for (int tries = 0; tries < 4; ++tries, hl = std::min(hl + 4, kPuzzleSlotBits)) {
        QueryDesc q;
        Plan p;
        if ((rc = puzzle_fill_desc(c, n, tiles, heaps, hl, q, p))) return rc;
        rc = run_wide(c, q, p);
        if (rc) return rc;
        kms += c->last_kernel_ms;
        if (c->last_r.status != ST_CAPACITY) break;
    }
    fill_result(c->last_q, c->last_r, res);
    res->kernel_ms = kms;
    if (res->status < 0) fail(c, res->status, "puzzle solve failed with status %d", res->status);
    return res->status;
}

extern "C" int ua_puzzle_solve_bounded(ua_ctx *c, int n, const uint8_t *tiles, uint64_t mem_bytes, ua_result *res) {
    if (!c || !res || !tiles) return fail(c, UA_ERR_INVALID, "null argument");
    c->last_valid = false;
    std::memset(res, 0, sizeof *res);
    if (!board_ok(n, tiles)) return fail(c, UA_ERR_INVALID, "board is not a permutation of 0..n*n-1");
    if (mem_bytes < (16ull << 20)) return fail(c, UA_ERR_INVALID, "memory bound below 16 MiB");
    if (!board_solvable(n, tiles)) {
        res->status = UA_NO_PATH;
        return UA_NO_PATH;
    }
    int rc = ensure_pdb(c, n);
    if (rc) return rc;
    // half the bound for the hash (keys and values), half for the heaps; a slot is 16
    // bytes (board, value) for N = 3, 4 and 20 for N = 5 (two board words, a value)
    const size_t slot_bytes = n == 5 ? 20 : 16;
    double kms = 0;  // every attempt's launch
    int hl = 10;
    while (hl < 32 && (slot_bytes << (hl + 1)) <= mem_bytes / 2) ++hl;
    for (int ws = 0; ws <= 3; ++ws) {
        // A round pops up to 64 keys a heap. The weighted searches run on fewer heaps, so
        // that the hash lasts some 500 rounds of about 3 new boards a pop: a narrower
        // frontier reaches deeper in the same memory.
        int heaps = plain_heaps(c, KIND_PUZZLE);
        if (ws > 0) {
            int64_t want = ((1ll << hl) / (64 * 3 * 512)) / kBlockWaves * kBlockWaves;
            heaps = (int)std::max<int64_t>(kBlockWaves, std::min<int64_t>(heaps, want));
        }
        const int64_t per_heap = (int64_t)(mem_bytes / 2 / 512) / heaps;  // 64-key nodes per heap
        if (per_heap < 3) return fail(c, UA_ERR_INVALID, "memory bound too small for %d heaps", heaps);
        QueryDesc q;
        Plan p;
        if ((rc = puzzle_fill_desc(c, n, tiles, heaps, hl, q, p))) return rc;
        q.heap_cap = (int)std::min<int64_t>(per_heap + 1, 1 << 30);
        p.heap_bytes = (size_t)heaps * (size_t)(q.heap_cap - 1) * 512;
        if (p.total() > mem_bytes)  // the plan's hash and heaps must keep to the bound
            return fail(c, UA_ERR_INVALID, "bounded plan of %zu bytes exceeds the bound of %llu", p.total(),
                        (unsigned long long)mem_bytes);
        q.wshift = ws;
        q.bounded = ws > 0;
        rc = run_wide(c, q, p);
        if (rc) return rc;
        kms += c->last_kernel_ms;
        const QueryResult &r = c->last_r;
        if (r.status == ST_OK && r.found) break;
        if (ws == 0 && r.status != ST_CAPACITY && r.status != ST_RANGE) break;  // no path, or an error
    }
    fill_result(c->last_q, c->last_r, res);
    res->kernel_ms = kms;
    if (res->status == UA_NO_PATH && c->last_q.bounded) res->status = UA_ERR_CAPACITY;  // pruned away
    if (res->status < 0) fail(c, res->status, "bounded puzzle solve failed with status %d", res->status);
    return res->status;
}

extern "C" int ua_puzzle_moves(ua_ctx *c, int32_t *dirs, int64_t cap, int64_t *len) {
    if (!c || !len) return fail(c, UA_ERR_INVALID, "null argument");
    if (!c->last_valid || c->last_q.kind != KIND_PUZZLE || !c->last_r.found)
        return fail(c, UA_ERR_INVALID, "no solved puzzle query");
    std::vector<int32_t> p;
    int rc = fetch_path(c, c->last_q, c->last_r, p);
    if (rc) return rc;
    return copy_path(c, p, dirs, cap, len);
}

// ------------------------------------------------------------------ batch

namespace uart {

// Waves per team of a batched / served query (a multiple of kBlockWaves: one block
// for grids by default, several for 15-puzzles), chosen per instance class -- puzzles,
// grids below and at or above grid_large_cells cells -- by the configuration or the
// autotuner.
int class_of(ua_ctx *c, int kind, int64_t cells) {
    if (kind == KIND_PUZZLE) return CLS_PUZZLE;
    return cells >= c->tune.grid_large_cells ? CLS_GRID_LARGE : CLS_GRID;
}

int team_waves(ua_ctx *c, int kind, int64_t cells) {
    int w;
    if (kind == KIND_PUZZLE) w = c->cfg.puzzle_team_waves > 0 ? c->cfg.puzzle_team_waves : c->tune.puzzle_waves;
    else if (c->cfg.team_waves > 0) w = c->cfg.team_waves;
    else w = class_of(c, kind, cells) == CLS_GRID_LARGE ? c->tune.grid_large_waves : c->tune.grid_waves;
    int per_cu = 1;
    if (teams_max_blocks(&per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
    int blocks = std::max(1, std::min((w + kBlockWaves - 1) / kBlockWaves, c->num_cu * per_cu));
    return blocks * kBlockWaves;
}

// XCD split of a batch or serving launch (see LaunchShape): the one-block classes get XCDs
// of their own, as many as their blocks need (at least one, and at least one is left to the
// other classes). Then each side's classes grow into their side's spare blocks (up to their
// query counts), or shrink, largest first and down to one group, until they fit it. The
// shape is left unsplit when one side is empty, the device's blocks are not a multiple of
// kXcds, or a side cannot fit.
static void split_xcds(LaunchShape &shape, int resident) {
    if (resident % kXcds) return;
    const int per = resident / kXcds;
    LaunchShape s = shape;
    int mask = 0, blocks[2] = {0, 0};  // [0] the other classes, [1] the split-off ones
    for (int k = 0; k < kClasses; ++k) {
        if (!s.nq[k]) continue;
        const int side = s.bpg[k] == 1;
        if (side) mask |= 1 << k;
        blocks[side] += s.groups[k] * s.bpg[k];
    }
    if (!blocks[0] || !blocks[1]) return;
    const int iso = std::min(kXcds - 1, std::max(1, (blocks[1] + per - 1) / per));
    const int cap[2] = {(kXcds - iso) * per, iso * per};
    for (int side = 0; side < 2; ++side) {
        auto on_side = [&](int k) { return s.nq[k] && (((mask >> k) & 1) == side); };
        while (blocks[side] > cap[side]) {
            int big = -1;
            for (int k = 0; k < kClasses; ++k)
                if (on_side(k) && s.groups[k] > 1 &&
                    (big < 0 || s.groups[k] * s.bpg[k] > s.groups[big] * s.bpg[big]))
                    big = k;
            if (big < 0) return;
            --s.groups[big];
            blocks[side] -= s.bpg[big];
        }
        for (int k = 0; k < kClasses; ++k)
            while (on_side(k) && s.groups[k] < s.nq[k] && blocks[side] + s.bpg[k] <= cap[side]) {
                ++s.groups[k];
                blocks[side] += s.bpg[k];
            }
    }
    s.iso_mask = (int16_t)mask;
    s.iso_xcds = (int16_t)iso;
    shape = s;
}

// Runs the planned queries of one problem set in one cooperative launch of k_teams.
// Each query's class (CLS_*) gets groups of heaps / kBlockWaves blocks; the resident
// blocks are shared among the classes in proportion to their demand (queries x blocks
// per team). Results come back in the order of descs.
static int run_set(ua_ctx *c, int set, std::vector<QueryDesc> &descs, std::vector<QueryResult> &res, int mode,
                   int64_t budget_ticks, const std::vector<QueryResult> *init) {
    const int nq = (int)descs.size();
    res.assign(nq, QueryResult{});
    if (!nq) return UA_OK;
    int per_cu = 0;
    HIPCHK(c, teams_max_blocks(&per_cu));
    const int resident = std::max(1, c->num_cu * std::max(1, per_cu));
    std::vector<int> order;
    order.reserve(nq);
    LaunchShape shape{};
    for (int k = 0; k < kClasses; ++k) {
        shape.bpg[k] = 0;
        shape.qoff[k] = (int)order.size();
        for (int i = 0; i < nq; ++i) {
            const QueryDesc &q = descs[i];
            if (q.cls != k) continue;
            int bpg = q.heaps / kBlockWaves;
            if (bpg < 1 || q.heaps % kBlockWaves || (shape.bpg[k] && shape.bpg[k] != bpg) || bpg > resident)
                return fail(c, UA_ERR_INVALID, "query %d planned for %d heaps, not a team width of its class", i,
                            q.heaps);
            shape.bpg[k] = bpg;
            order.push_back(i);
            ++shape.nq[k];
        }
        if (!shape.bpg[k]) shape.bpg[k] = 1;
        shape.segcap[k] = shape.bpg[k] * kBlockWaves * 64 * 8;  // at most 8 successors per pop
    }
    // widened queries first (a group each), then the rest in proportion to demand
    int used = 0;
    if (shape.nq[CLS_WIDE]) {
        shape.groups[CLS_WIDE] = std::max(1, std::min(shape.nq[CLS_WIDE], resident / 2 / shape.bpg[CLS_WIDE]));
        used = shape.groups[CLS_WIDE] * shape.bpg[CLS_WIDE];
    }
    double demand = 0;
    for (int k = 0; k < kClasses; ++k)
        if (k != CLS_WIDE) demand += (double)shape.nq[k] * shape.bpg[k];
    const int rest = resident - used;
    for (int k = 0; k < kClasses; ++k) {
        if (!shape.nq[k] || k == CLS_WIDE) continue;
        int want = (int)(rest * shape.nq[k] * shape.bpg[k] / demand);
        shape.groups[k] = std::max(1, std::min(shape.nq[k], want / shape.bpg[k]));
        used += shape.groups[k] * shape.bpg[k];
    }
    while (used > resident) {  // rounding up to one group per class can overshoot
        int big = 0;
        for (int k = 1; k < kClasses; ++k)
            if (shape.groups[k] * shape.bpg[k] > shape.groups[big] * shape.bpg[big]) big = k;
        if (shape.groups[big] <= 1) return fail(c, UA_ERR_INVALID, "launch does not fit the device");
        --shape.groups[big];
        used -= shape.bpg[big];
    }
    if (mode != 0 && !(c->cfg.flags & UA_FLAG_XCD_SHARED)) split_xcds(shape, resident);
    shape.budget_ticks = budget_ticks;
    std::vector<QueryDesc> sorted(nq);
    for (int k = 0; k < nq; ++k) sorted[k] = descs[order[k]];
    std::vector<QueryResult> sinit;
    if (init) {
        sinit.resize(nq);
        for (int k = 0; k < nq; ++k) sinit[k] = (*init)[order[k]];
    }
    // every resident block: the grid of a launch stays the same from step to step
    std::vector<QueryResult> sres(nq);
    int rc = run_step(c, mode, set, resident, shape, sorted.data(), init ? sinit.data() : nullptr, nq, sres.data());
    if (rc) return rc;
    for (int k = 0; k < nq; ++k) res[order[k]] = sres[k];
    return UA_OK;
}

// The planned queries in one launch per problem set present (5x5 puzzles run in a
// launch of their own after the rest); last_kernel_ms is the launches' sum.
int run_batch(ua_ctx *c, std::vector<QueryDesc> &descs, std::vector<QueryResult> &res, int mode,
              int64_t budget_ticks, const std::vector<QueryResult> *init) {
    std::vector<int> idx[kSets];
    for (int i = 0; i < (int)descs.size(); ++i) idx[set_of(descs[i])].push_back(i);
    int present = 0, only = SET_MAIN;
    for (int set = 0; set < kSets; ++set)
        if (!idx[set].empty()) ++present, only = set;
    if (present <= 1) return run_set(c, only, descs, res, mode, budget_ticks, init);
    res.assign(descs.size(), QueryResult{});
    double kms = 0;
    for (int set = 0; set < kSets; ++set) {
        if (idx[set].empty()) continue;
        std::vector<QueryDesc> d;
        std::vector<QueryResult> r, in;
        for (int i : idx[set]) {
            d.push_back(descs[i]);
            if (init) in.push_back((*init)[i]);
        }
        int rc = run_set(c, set, d, r, mode, budget_ticks, init ? &in : nullptr);
        if (rc) return rc;
        kms += c->last_kernel_ms;
        for (size_t k = 0; k < idx[set].size(); ++k) res[idx[set][k]] = r[k];
    }
    c->last_kernel_ms = kms;
    return UA_OK;
}

}  // namespace uart

extern "C" int ua_batch_solve(ua_ctx *c, const ua_query *qs, int64_t nq, ua_result *out) {
    if (!c || (!qs && nq) || (!out && nq) || nq < 0) return fail(c, UA_ERR_INVALID, "null argument");
    // The previous batch's state goes; its largest arena stays as a spare for this batch's
    // first sub-batch that fits in it (allocating gigabytes per call costs tens of ms).
    DevBuf spare;
    struct Release {  // the spare is freed on every way out, errors included
        DevBuf &b;
        ~Release() { b.release(); }
    } spare_guard{spare};
    for (auto &b : c->b_arenas) {
        if (b.n > spare.n) std::swap(spare, b);
        b.release();
    }
    c->b_arenas.clear();
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
    size_t budget = c->cfg.mem_budget ? c->cfg.mem_budget : (size_t)((free_b + spare.n) * 0.85);
    // a spare arena near the budget sets it, so that the sub-batches reuse the allocation
    // as it is instead of freeing it for a slightly larger one
    if (!c->cfg.mem_budget && spare.n < budget && spare.n >= budget / 10 * 9) budget = spare.n;

    // plan every query; reject bad ones up front
    std::vector<QueryDesc> descs(nq);
    std::vector<Plan> plans(nq);
    std::vector<int32_t> pre(nq, 1);  // 1 = to run, else final status
    for (int64_t i = 0; i < nq; ++i) {
        const ua_query &u = qs[i];
        int rc;
        if (u.kind == UA_QUERY_GRID) {
            rc = grid_fill_desc(c, u.grid, u.sx, u.sy, u.ex, u.ey,
                                team_waves(c, KIND_GRID, u.grid ? (int64_t)u.grid->H * u.grid->W : 0), 1, descs[i],
                                plans[i]);
        } else if (u.kind == UA_QUERY_PUZZLE) {
            if (!board_ok(u.n, u.tiles)) rc = fail(c, UA_ERR_INVALID, "query %lld: bad board", (long long)i);
            else if (!board_solvable(u.n, u.tiles)) rc = UA_NO_PATH;
            else if ((rc = ensure_pdb(c, u.n)) == UA_OK)
                rc = puzzle_fill_desc(c, u.n, u.tiles, team_waves(c, KIND_PUZZLE, 0), batch_hash_log2(c, u.n),
                                      descs[i], plans[i]);
        } else {
            rc = fail(c, UA_ERR_INVALID, "query %lld: bad kind", (long long)i);
        }
        if (rc) pre[i] = rc;
        else if (plans[i].total() > budget) pre[i] = fail(c, UA_ERR_OOM, "query %lld exceeds the memory budget", (long long)i);
        else pre[i] = 1;
    }
    std::vector<QueryResult> results(nq);
    for (int64_t i = 0; i < nq; ++i) {
        std::memset(&out[i], 0, sizeof out[i]);
        if (pre[i] != 1) out[i].status = pre[i];
    }

    // largest estimated work first (longest-processing-time order), then admit in
    // memory-bounded rounds; queries that overflow their heap/hash retry bigger.
    std::vector<int64_t> todo;
    for (int64_t i = 0; i < nq; ++i)
        if (pre[i] == 1) todo.push_back(i);
    std::stable_sort(todo.begin(), todo.end(), [&](int64_t a, int64_t b) { return plans[a].total() > plans[b].total(); });
    std::vector<int> scale(nq, 1);
    c->b_desc.assign(nq, QueryDesc{});
    c->b_res.assign(nq, QueryResult{});
    c->b_arena_of.assign(nq, -1);
    c->b_paths.assign(nq, {});
    c->b_kept.assign(nq, 0);
    size_t held = 0;
    for (int attempt = 0; attempt < 4 && !todo.empty(); ++attempt) {
        std::vector<int64_t> retry;
        size_t pos = 0;
        while (pos < todo.size()) {
            // admit a prefix of todo that fits in the budget
            auto need = [&](int64_t i) {
                return align256(plans[i].best_bytes) + align256(plans[i].vals_bytes) + align256(plans[i].heap_bytes);
            };
            size_t used = 0, end = pos;
            while (end < todo.size() && used + need(todo[end]) <= budget) used += need(todo[end++]);
            if (end == pos) {
                out[todo[pos]].status = fail(c, UA_ERR_OOM, "query does not fit the budget");
                pre[todo[pos]] = UA_ERR_OOM;
                ++pos;
                continue;
            }
            if (held + used > budget) {
                // release earlier sub-batches' state; their paths are walked first and kept
                std::vector<const QueryDesc *> rq;
                std::vector<const QueryResult *> rr;
                std::vector<int64_t> ri;
                for (int64_t i = 0; i < nq; ++i)
                    if (c->b_arena_of[i] >= 0 && c->b_res[i].found && c->b_res[i].status == ST_OK) {
                        rq.push_back(&c->b_desc[i]);
                        rr.push_back(&c->b_res[i]);
                        ri.push_back(i);
                    }
                std::vector<std::vector<int32_t>> kept(ri.size());
                int prc = fetch_paths(c, (int)ri.size(), rq.data(), rr.data(), kept.data());
                if (prc) return prc;
                for (size_t k = 0; k < ri.size(); ++k) {
                    c->b_paths[ri[k]] = std::move(kept[k]);
                    c->b_kept[ri[k]] = 1;
                }
                // the largest arena stays as the spare of the next sub-batch: allocating and
                // freeing hundreds of gigabytes costs seconds (about 40 GB/s measured)
                for (auto &b : c->b_arenas) {
                    if (b.n > spare.n) std::swap(spare, b);
                    b.release();
                }
                for (auto &a : c->b_arena_of) a = -1;
                held = 0;
            }
            c->b_arenas.emplace_back();
            DevBuf &arena = c->b_arenas.back();
            if (spare.p && spare.n >= used) {
                std::swap(arena, spare);
            } else {
                spare.release();
                // when more sub-batches follow, the arena takes the whole budget so that each
                // of them can reuse it
                size_t rest = 0;
                for (size_t k = end; k < todo.size(); ++k) rest += need(todo[k]);
                HIPCHK(c, arena.ensure(rest ? std::max(used, std::min(budget, used + rest)) : used));
            }
            held += arena.n;
            int arena_idx = (int)c->b_arenas.size() - 1;
            char *base = arena.as<char>();
            // layout: [grid best tables] (written by the init kernel), [5x5 puzzle vals] filled
            // with 0xFF, [puzzle slots] = 0, then heaps
            std::vector<QueryDesc> sub;
            size_t off = 0;
            for (size_t k = pos; k < end; ++k) {
                int64_t i = todo[k];
                QueryDesc q = descs[i];
                if (q.kind == KIND_GRID) {
                    q.best = reinterpret_cast<uint64_t *>(base + off);
                    off += align256(plans[i].best_bytes);
                }
                sub.push_back(q);
            }
            const size_t ff_begin = off;
            for (size_t k = pos; k < end; ++k) {
                int64_t i = todo[k];
                if (descs[i].kind == KIND_PUZZLE) {
                    sub[k - pos].vals = reinterpret_cast<uint32_t *>(base + off);
                    off += align256(plans[i].vals_bytes);
                }
            }
            size_t ff_end = off;
            for (size_t k = pos; k < end; ++k) {
                int64_t i = todo[k];
                if (descs[i].kind == KIND_PUZZLE) {
                    sub[k - pos].best = reinterpret_cast<uint64_t *>(base + off);
                    off += align256(plans[i].best_bytes);
                }
            }
            size_t zero_end = off;
            for (size_t k = pos; k < end; ++k) {
                sub[k - pos].heap = reinterpret_cast<uint64_t *>(base + off);
                off += align256(plans[todo[k]].heap_bytes);
            }
            for (const QueryDesc &q : sub)
                if (q.kind == KIND_GRID) HIPCHK(c, launch_grid_init_best(q.best, q.bitmap, q.H, q.W, q.wpr, c->st));
            if (ff_end > ff_begin) HIPCHK(c, hipMemsetAsync(base + ff_begin, 0xFF, ff_end - ff_begin, c->st));
            if (zero_end > ff_end) HIPCHK(c, hipMemsetAsync(base + ff_end, 0, zero_end - ff_end, c->st));
            std::vector<QueryResult> sr;
            int rc = run_batch(c, sub, sr, 1);
            if (rc) return rc;
            for (size_t k = pos; k < end; ++k) {
                int64_t i = todo[k];
                c->b_desc[i] = sub[k - pos];
                c->b_res[i] = sr[k - pos];
                c->b_arena_of[i] = arena_idx;
                if (sr[k - pos].status == ST_CAPACITY && attempt < 3) retry.push_back(i);
            }
            pos = end;
        }
        for (int64_t i : retry) {  // replan bigger: grids 4x the open list, puzzles 16x the hash
            scale[i] *= 4;
            if (descs[i].kind == KIND_GRID) {
                const ua_query &u = qs[i];
                grid_fill_desc(c, u.grid, u.sx, u.sy, u.ex, u.ey, descs[i].heaps, scale[i], descs[i],
                               plans[i]);
            } else {
                // a puzzle that filled its hash is a heavy one: 16x at once spares the second
                // retry that 4x often needed (each retry reruns the search from its start),
                // and a team 4x as wide, up to a quarter of the device, since a retry launch
                // holds few queries
                int hl = std::min(descs[i].hash_log2 + 4, kPuzzleSlotBits);
                int per_cu = 1;
                if (teams_max_blocks(&per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
                const int quarter = std::max(kBlockWaves, c->num_cu * per_cu * kBlockWaves / 4 / kBlockWaves * kBlockWaves);
                const int heaps = std::max(descs[i].heaps, std::min(4 * descs[i].heaps, quarter));
                const int cls = descs[i].cls;
                puzzle_fill_desc(c, qs[i].n, qs[i].tiles, heaps, hl, descs[i], plans[i]);
                descs[i].cls = cls;
            }
        }
        todo.swap(retry);
    }
    for (int64_t i = 0; i < nq; ++i) {
        if (pre[i] != 1 || c->b_desc[i].kind == 0) continue;
        fill_result(c->b_desc[i], c->b_res[i], &out[i]);
        out[i].latency_ms = out[i].device_ms;
    }
    return UA_OK;
}

extern "C" int ua_batch_path(ua_ctx *c, int64_t i, int32_t *outp, int64_t cap, int64_t *len) {
    if (!c || !len || i < 0 || i >= (int64_t)c->b_desc.size()) return fail(c, UA_ERR_INVALID, "bad query index");
    if (c->b_arena_of[i] < 0) {  // state released by a later sub-batch: the path walked before
        if (!c->b_kept[i]) return fail(c, UA_ERR_INVALID, "query %lld has no path", (long long)i);
        return copy_path(c, c->b_paths[i], outp, cap, len);
    }
    std::vector<int32_t> p;
    int rc = fetch_path(c, c->b_desc[i], c->b_res[i], p);
    if (rc) return rc;
    return copy_path(c, p, outp, cap, len);
}

// ------------------------------------------------------------------ diagnostics

extern "C" int ua_debug_heap_selftest(ua_ctx *c, const int32_t *ops, int nops, const uint64_t *keys, int64_t nkeys,
                                      int cap_nodes, uint64_t *out_pops, int64_t out_cap, int32_t *final_size) {
    if (!c || !ops || nops <= 0 || cap_nodes < 2) return fail(c, UA_ERR_INVALID, "bad arguments");
    int64_t need_keys = 0, npops = 0;
    for (int i = 0; i < nops; ++i) {
        if (ops[i] > 64) return fail(c, UA_ERR_INVALID, "insert batches hold at most 64 keys");
        if (ops[i] >= 0) need_keys += ops[i];
        else ++npops;
    }
    if (need_keys > nkeys || npops * 64 > out_cap) return fail(c, UA_ERR_INVALID, "buffers too small");
    DevBuf d_ops, d_keys, d_store, d_out, d_st;
    HIPCHK(c, d_ops.ensure(sizeof(int) * nops));
    HIPCHK(c, d_keys.ensure(8 * std::max<int64_t>(need_keys, 1)));
    HIPCHK(c, d_store.ensure((size_t)(cap_nodes - 1) * 512));
    HIPCHK(c, d_out.ensure(8 * std::max<int64_t>(npops * 64, 1)));
    HIPCHK(c, d_st.ensure(8));
    HIPCHK(c, hipMemcpy(d_ops.p, ops, sizeof(int) * nops, hipMemcpyHostToDevice));
    if (need_keys) HIPCHK(c, hipMemcpy(d_keys.p, keys, 8 * need_keys, hipMemcpyHostToDevice));
    HIPCHK(c, launch_heap_selftest(d_ops.as<int>(), nops, d_keys.as<uint64_t>(), d_store.as<uint64_t>(), cap_nodes,
                                   d_out.as<uint64_t>(), d_st.as<int>(), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (npops) HIPCHK(c, hipMemcpy(out_pops, d_out.p, 8 * npops * 64, hipMemcpyDeviceToHost));
    int st = 0;
    HIPCHK(c, hipMemcpy(&st, d_st.p, 4, hipMemcpyDeviceToHost));
    if (final_size) *final_size = st;
    for (DevBuf *b : {&d_ops, &d_keys, &d_store, &d_out, &d_st}) b->release();
    return UA_OK;
}

// Random device-memory atomic throughput (k_atomic_probe): blocks x 256 threads x iters
// atomicMins over 2^log2_words 32-bit words, timed with HIP events; *gops = atomics / ns.
extern "C" int ua_debug_atomic_rate(ua_ctx *c, int log2_words, int blocks, int iters, double *gops) {
    if (!c || !gops || log2_words < 10 || log2_words > 32 || blocks < 1 || iters < 1)
        return fail(c, UA_ERR_INVALID, "bad arguments");
    const uint64_t n = 1ull << log2_words;
    DevBuf a, sink;
    HIPCHK(c, a.ensure(n * 4));
    HIPCHK(c, sink.ensure(64));
    HIPCHK(c, hipMemsetAsync(a.p, 0xFF, n * 4, c->st));
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up; the second is timed
        HIPCHK(c, hipEventRecord(c->ev0, c->st));
        HIPCHK(c, launch_atomic_probe(a.as<uint32_t>(), n - 1, blocks, iters, sink.as<uint32_t>(), c->st));
        HIPCHK(c, hipEventRecord(c->ev1, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    }
    a.release();
    sink.release();
    *gops = (double)blocks * 256.0 * iters / ((double)ms * 1e6);
    return UA_OK;
}

// Dependent-access latency (k_latency_probe): one wave follows a random cycle over
// 2^log2_words 64-bit words (built on the host, every entry the next index), `iters`
// steps of access `kind` (0 workgroup load, 1 workgroup atomicMin, 2 agent atomicMin,
// 3 agent load); *ns = nanoseconds per step of the second of two runs.
extern "C" int ua_debug_access_latency(ua_ctx *c, int log2_words, int kind, int iters, double *ns) {
    if (!c || !ns || log2_words < 6 || log2_words > 28 || kind < 0 || kind > 3 || iters < 1)
        return fail(c, UA_ERR_INVALID, "bad arguments");
    const uint64_t n = 1ull << log2_words;
    // one random cycle through all n entries (Sattolo's shuffle of the identity, driven by
    // Marsaglia's xorshift64), so that a chase never settles into a short, cached loop
    std::vector<uint64_t> h(n);
    for (uint64_t i = 0; i < n; ++i) h[i] = i;
    uint64_t x = 88172645463325252ull;
    for (uint64_t i = n - 1; i > 0; --i) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        std::swap(h[i], h[x % i]);
    }
    DevBuf a, out;
    HIPCHK(c, a.ensure(n * 8));
    HIPCHK(c, out.ensure(64));
    HIPCHK(c, hipMemcpy(a.p, h.data(), n * 8, hipMemcpyHostToDevice));
    uint64_t r[2] = {0, 0};
    for (int rep = 0; rep < 2; ++rep) {
        HIPCHK(c, launch_latency_probe(a.as<uint64_t>(), n - 1, kind, iters, out.as<uint64_t>(), c->st));
        HIPCHK(c, hipMemcpyAsync(r, out.p, 16, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    a.release();
    out.release();
    *ns = (double)r[0] * 10.0 / iters;
    return UA_OK;
}
