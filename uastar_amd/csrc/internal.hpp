// internal.hpp -- host runtime internals shared by runtime.cpp (plain and batch
// solves) and serve.cpp (the serving engine). Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/uastar.h"
#include "common.hpp"
#include "kernels.hpp"

namespace uart {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

inline int bits_for(int64_t v) {  // bits to hold 0..v-1 (at least 1)
    int b = 1;
    while ((1ll << b) < v) ++b;
    return b;
}

inline double ticks_ms(int64_t t0, int64_t t1) { return (double)(t1 - t0) / 1e5; }  // s_memrealtime: 100 MHz

constexpr double SQRT2 = 1.4142135623730951;

}  // namespace uart

struct ua_grid {
    int H, W, wpr;
    uint64_t *bm;
};

struct ServeEngine;
void ua_serve_release(ua_ctx *c);  // serve.cpp

namespace uart {
// The knobs the autotuner sets (tune.cpp). A nonzero ua_config field overrides the
// matching knob; the defaults are the untuned starting point.
struct Tuning {
    int grid_waves = 16;          // waves (= heaps) per batched / served grid query below grid_large_cells
    int grid_large_waves = 16;    // ... and at or above grid_large_cells
    int64_t grid_large_cells = 1ll << 22;
    int grid_bucket_window = 160;  // buckets (1/16 step of f each) a bucket-ring round takes keys from
    int puzzle_waves = 64;        // waves (= heaps) per batched / served puzzle query
    int puzzle_hash_log2[6] = {0, 0, 0, 17, 24, 22};   // batched / served puzzles, by n
    int plain_heaps[2] = {0, 0};  // plain whole-GPU solve, by kind (grid, puzzle); 0 = 4 per CU
    int plain_hash_log2[6] = {0, 0, 0, 19, 24, 24};    // plain puzzles, by n
    bool tuned = false;
};
}  // namespace uart

namespace uart {
// An instantiated graph of one k_teams launch (step.cpp): H2D copy of the I/O block,
// timing event, the cooperative search kernel, timing event, D2H copy of the results.
// Rebuilt when an address or the block count it was built for changes.
struct StepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    hipGraphNode_t h2d = nullptr, kern = nullptr, d2h = nullptr;
    void *io = nullptr, *host = nullptr, *segs = nullptr;
    int mode = -1, set = 0, blocks = 0;
    void release();
};
}  // namespace uart

struct ua_ctx {
    int dev = 0;
    hipStream_t st = nullptr;
    // device address of this GPU's copy of the allocation-free entry point's scratch array
    // (runtime.cpp), from hipGetSymbolAddress at ua_create; null if that call failed
    char *noalloc_scratch = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around each search launch on st
    double last_kernel_ms = 0;
    ua_config cfg{};
    uart::Tuning tune;
    std::string err;
    int num_cu = 0;
    size_t total_mem = 0;
    uart::DevBuf pdb[6];  // by side n (3, 4, 5)
    uart::DevBuf d_added;
    // launches of k_teams (step.cpp): the I/O block, the groups' insert buffers, the
    // pinned staging copy of the I/O block and one graph per mode
    uart::DevBuf io, segs;
    char *h_io = nullptr;
    size_t h_io_n = 0;
    uart::StepGraph graph[8];        // by (mode, problem set, block count); reused round robin
    int graph_next = 0;
    bool graphs = false;             // replay launches from graphs (UA_FLAG_GRAPH)
    std::string graph_err;           // why graphs were switched off, if they were
    int64_t graph_launches = 0, direct_launches = 0;
    double step_wall_ms = 0, step_kernel_ms = 0;
    // plain solve
    uart::DevBuf best, vals, heap, pathbuf;
    ua::QueryDesc last_q{};
    ua::QueryResult last_r{};
    bool last_valid = false;
    // batch
    std::vector<uart::DevBuf> b_arenas;        // one per admitted sub-batch, kept for ua_batch_path
    std::vector<int> b_arena_of;         // arena index of each query, -1 = state released
    std::vector<ua::QueryDesc> b_desc;
    std::vector<ua::QueryResult> b_res;
    std::vector<int64_t> b_order;
    std::vector<std::vector<int32_t>> b_paths;  // paths walked before their sub-batch's state was released
    std::vector<char> b_kept;
    // serving
    std::mutex mu;
    ServeEngine *serve = nullptr;
};


namespace uart {
using namespace ua;

int fail(ua_ctx *c, int code, const char *fmt, ...);

#define HIPCHK(c, expr)                                                                          \
    do {                                                                                         \
        hipError_t e__ = (expr);                                                                 \
        if (e__ != hipSuccess)                                                                   \
            return uart::fail(c, e__ == hipErrorOutOfMemory ? UA_ERR_OOM : UA_ERR_HIP,            \
                              "%s: %s (%s:%d)", #expr, hipGetErrorString(e__), __FILE__, __LINE__); \
    } while (0)

struct Plan {
    size_t best_bytes = 0, vals_bytes = 0, heap_bytes = 0;
    size_t total() const { return best_bytes + vals_bytes + heap_bytes; }
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int status_of(const QueryResult &r);
int grid_fill_desc(ua_ctx *c, const ua_grid *g, int sx, int sy, int ex, int ey, int heaps, int heap_scale,
                   QueryDesc &q, Plan &p);
bool board_ok(int n, const uint8_t *t);
bool board_solvable(int n, const uint8_t *t);
uint64_t pack_board(int n, const uint8_t *t);
void pack_board5(const uint8_t *t, uint64_t *w0, uint64_t *w1);
int puzzle_fill_desc(ua_ctx *c, int n, const uint8_t *tiles, int heaps, int hash_log2, QueryDesc &q, Plan &p);
int ensure_pdb(ua_ctx *c, int n);
void fill_result(const QueryDesc &q, const QueryResult &r, ua_result *o);
// start..goal cell ids (grid) or blank moves (puzzle) of a finished query whose state is
// still on the device
int fetch_path(ua_ctx *c, const QueryDesc &q, const QueryResult &r, std::vector<int32_t> &out);
// the same for n queries in one launch of the walk kernel (out[i] empty: no path)
int fetch_paths(ua_ctx *c, int n, const QueryDesc *const *qs, const QueryResult *const *rs,
                std::vector<int32_t> *out);
// waves per team (a multiple of kBlockWaves) of a batched / served query of this kind
// and size, and its instance class (CLS_*)
int team_waves(ua_ctx *c, int kind, int64_t cells);
int class_of(ua_ctx *c, int kind, int64_t cells);
int plain_heaps(ua_ctx *c, int kind, int64_t cells = 0);
int batch_hash_log2(ua_ctx *c, int n);
int plain_hash_log2(ua_ctx *c, int n);
// One k_teams launch of `blocks` blocks for problem set `set` (step.cpp): the shape, the nq descriptors and
// the results reported for queries the launch does not reach (init, or zeros) go to the
// device in one copy; the results come back into out. Sets c->last_kernel_ms.
int run_step(ua_ctx *c, int mode, int set, int blocks, const LaunchShape &shape, const QueryDesc *descs,
             const QueryResult *init, int nq, QueryResult *out);
void release_steps(ua_ctx *c);
// mode: 1 = a ua_batch_solve sub-batch, 2 = a serving step (names the kernel in profiles).
// budget_ticks > 0: each group stops after that much device time (serving); `init`, if
// given, is the result reported for a query the launch does not reach.
int run_batch(ua_ctx *c, std::vector<QueryDesc> &descs, std::vector<QueryResult> &res, int mode,
              int64_t budget_ticks = 0, const std::vector<QueryResult> *init = nullptr);

}  // namespace uart
