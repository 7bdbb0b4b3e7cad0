/*
 * uastar.h -- C ABI of libuastar, the MI355X (gfx950) GA* engine.
 *
 * Plain pointers and sizes only. Every entry point returns a status code
 * (UA_OK = 0, UA_NO_PATH = 1, negative = error; see ua_last_error()).
 * Citations name the uAstar interface each entry point replaces (paths relative
 * to the uAstar source tree).
 *
 * Grid bitmaps: row-padded, 1 bit per cell, 1 = open. Row x occupies words
 * [x*wpr, (x+1)*wpr) with wpr = (W + 63) / 64; cell (x, y) is bit (y & 63) of word
 * x*wpr + (y >> 6). Cell ids are x*W + y (pathway.hpp:99-108). A move u->v is legal
 * when u is open and v is in range; a blocked v is a dead end and may only be the
 * goal (pathway/CPU-solver.cpp:64-70 with all-or-nothing cell masks).
 * Costs are exact: orth + diag*sqrt(2) with integer step counts.
 */
#ifndef UASTAR_H
#define UASTAR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define UA_OK 0
#define UA_NO_PATH 1          /* search exhausted without reaching the goal */
#define UA_ERR_INVALID (-1)   /* bad argument */
#define UA_ERR_OOM (-2)       /* device allocation failed / over the memory budget */
#define UA_ERR_CAPACITY (-3)  /* heap or hash capacity exhausted (retries exhausted) */
#define UA_ERR_RANGE (-4)     /* cost above the exact-key limit (180000) */
#define UA_ERR_HIP (-5)       /* HIP runtime error */
#define UA_ERR_NO_DEVICE (-6) /* no usable GPU */
#define UA_ERR_ROUNDS (-7)    /* round limit reached */
#define UA_ERR_NO_PDB (-8)    /* puzzle solve before ua_puzzle_build_pdb */
#define UA_ERR_ABORT (-9)     /* a device barrier timed out; results discarded */

#define UA_QUERY_GRID 1
#define UA_QUERY_PUZZLE 2

typedef struct ua_ctx ua_ctx;
typedef struct ua_grid ua_grid; /* a device-resident grid bitmap */

typedef struct {
    int32_t team_waves;        /* waves (= heaps) per grid query in batch/serve (multiple of 16); 0 = tuned */
    int32_t plain_heaps;       /* heaps (= waves) of the plain whole-GPU solve; 0 = tuned (untuned: 16 = one
                                  block on the bucket ring for grids below 2^22 cells, 256 for larger
                                  grids, 4 per CU for puzzles) */
    int32_t max_rounds;        /* 0 = 1<<26 */
    int32_t hash_log2;         /* puzzle hash table size per query; 0 = tuned */
    uint64_t mem_budget;       /* device bytes the batch/serve planner may use; 0 = 80-85% of free */
    int32_t puzzle_team_waves; /* waves (= heaps) per puzzle query in batch/serve (multiple of 16); 0 = tuned */
    int32_t flags;             /* UA_FLAG_* bits; 0 = defaults */
} ua_config;

/* ua_config.flags: replay each search launch (copy in, cooperative kernel, copy out)
 * from an instantiated hipGraph instead of issuing it call by call */
#define UA_FLAG_GRAPH 1
/* ua_config.flags: a served query that finds its goal keeps its path (grid cell ids) or
 * moves (puzzle) on the host until ua_serve_path fetches them */
#define UA_FLAG_SERVE_PATHS 2
/* ua_config.flags: grids on one-block teams keep one sorted heap per wave instead of the
 * team's bucket ring (for comparisons; multi-block grid teams always use heaps) */
#define UA_FLAG_GRID_HEAPS 4
/* ua_config.flags: batch and serving launches spread every class over every XCD instead of
 * giving one-block grid teams XCDs (and so L2 caches) apart from the multi-block puzzle
 * teams (for comparisons) */
#define UA_FLAG_XCD_SHARED 8

typedef struct {
    int32_t status;
    int32_t found;
    int64_t orth, diag;  /* grid: exact cost = orth + diag*sqrt(2) */
    int64_t steps;       /* puzzle: optimal number of moves */
    double cost;         /* grid: orth + diag*sqrt(2) as a double */
    int64_t rounds, pops, expanded, generated, inserted;
    double device_ms;    /* search time on the device (s_memrealtime stamps of the team) */
    double kernel_ms;    /* duration of the search launch, from HIP events on the context's stream
                            (a plain puzzle solve that reran with a larger hash, or a bounded
                            solve that reran weighted: every attempt's launch; a batch: the
                            launch that ran this query, shared by its sub-batch) */
    double latency_ms;   /* serve: completion - submission; batch: completion - launch */
    int32_t optimal;     /* 1: the cost is proven optimal (every exact search that finds a path) */
    int32_t weight;      /* weight on h of the search that found the answer (1 = plain A*) */
    int64_t atomics;     /* device-memory atomics the search issued on its table or hash (relaxing
                            atomicMin, slot-claiming CAS) */
    int32_t heaps;       /* waves of the team that searched the query (= its heaps; a one-block grid
                            team shares one bucket ring) */
    int32_t reserved;
} ua_result;

/* Device context: owns every device allocation. Replaces the device set-up of
 * GPUPathwaySolver::initialize (pathway/GPU-solver.cu:91-140) and
 * GPUPuzzleSolver<N>::initialize (puzzle/GPU-solver.cuh:64-195); no device reset. */
int ua_create(int ordinal, const ua_config *cfg, ua_ctx **out);
void ua_destroy(ua_ctx *ctx);
const char *ua_last_error(ua_ctx *ctx);
/* 1 if a GPU is visible, without creating a context */
int ua_device_count(void);

/* ---- grids. Replace the H2D copy of Pathway::graph() (pathway/GPU-solver.cu:103). */
int ua_grid_create(ua_ctx *ctx, int H, int W, const uint64_t *host_bitmap, ua_grid **out);
/* From the byte-per-cell graph of Pathway::graph() (pathway.hpp:116-119): 0 = blocked,
 * 0xFF = open. A byte is a cell's out-move mask (bit d gates move d,
 * pathway/CPU-solver.cpp:64-66); every input module writes all-or-nothing masks
 * (input/custom.cpp:23, input/zigzag.cpp:26-39), and any other byte is rejected with
 * UA_ERR_INVALID (the grid is not created). */
int ua_grid_from_bytes(ua_ctx *ctx, int H, int W, const uint8_t *bytes, ua_grid **out);
/* i.i.d. Bernoulli(block_rate) obstacles from seed, generated on the device;
 * start and goal forced open (the advertised `random` input module, main.cpp:28-29 and
 * :129-130, which the reference never built) */
int ua_grid_generate(ua_ctx *ctx, int H, int W, uint64_t seed, double block_rate, int sx, int sy,
                     int ex, int ey, ua_grid **out);
int ua_grid_download(ua_ctx *ctx, const ua_grid *g, uint64_t *host_bitmap);
void ua_grid_free(ua_ctx *ctx, ua_grid *g);

/* Plain single-query solve on the whole GPU: replaces GPUPathwaySolver::solve
 * (pathway/GPU-solver.cu:142-290) -> UA_OK (found) or UA_NO_PATH. */
int ua_grid_solve(ua_ctx *ctx, const ua_grid *g, int sx, int sy, int ex, int ey, ua_result *res);
/* Path of the last ua_grid_solve, start..goal cell ids: replaces
 * GPUPathwaySolver::getSolution (pathway/GPU-solver.cu:292-316). */
int ua_grid_path(ua_ctx *ctx, int32_t *ids, int64_t cap, int64_t *len);

/* ua_grid_solve and ua_grid_path in one call that allocates nothing: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Every device buffer of the call is
 * a region of one static array of 46,158,080 bytes (44 MiB) per GPU, held until the process
 * ends; the grid may have
 * at most 2^20 cells (UA_ERR_INVALID otherwise, before any device work). The solve runs one
 * block on the bucket ring; a ring bucket that fills up returns UA_ERR_CAPACITY (no retry).
 * With len != NULL the path (start..goal cell ids) goes to path[0..*len); path == NULL asks
 * for the length only. Calls that overlap on one GPU share the array and race. */
int ua_grid_solve_noalloc(ua_ctx *ctx, const ua_grid *g, int sx, int sy, int ex, int ey, ua_result *res,
                          int32_t *path, int64_t cap, int64_t *len);

/* ---- sliding-tile puzzles, n = 3, 4 or 5. */
/* Builds the reference's additive disjoint PDB on the device (puzzle/database.cpp:37-84;
 * partitions of puzzle/CPU-solver.hpp:44-104: one 8-tile table for n = 3, 8 + 7 tiles for
 * n = 4, four 6-tile tables for n = 5); replaces the PDB upload of
 * GPUPuzzleSolver<N>::initialize (puzzle/GPU-solver.cuh:130-160). */
int ua_puzzle_build_pdb(ua_ctx *ctx, int n);
/* Copies the device PDB (all partitions concatenated) to the host. */
int ua_puzzle_download_pdb(ua_ctx *ctx, int n, uint8_t *host, uint64_t size);
/* Uploads a PDB (all partitions concatenated, e.g. read from uAstar's database_*.bin
 * cache files, puzzle/database.cpp:86-122) instead of building it; call before solving. */
int ua_puzzle_load_pdb(ua_ctx *ctx, int n, const uint8_t *host, uint64_t size);
/* Plain solve of one board (row-major, 0 = blank): replaces GPUPuzzleSolver<N>::solve
 * (puzzle/GPU-solver.cuh:197-274). Unsolvable boards return UA_NO_PATH at once. */
int ua_puzzle_solve(ua_ctx *ctx, int n, const uint8_t *tiles, ua_result *res);
/* The memory-bounded scheme the reference describes for puzzles too large for plain A*
 * (README.md:30-34) but never built. The search's hash and heaps get at most mem_bytes
 * of device memory. It first runs plain A*; if that does not fit, it runs weighted
 * searches (keys (g >> k) + h, k = 1, 2, 3: weight 2, 4, 8 on h) on fewer heaps in the
 * same memory, dropping successors that do not fit, and returns the first solution
 * found. Then
 * res->optimal = 0 and res->weight gives the weight. UA_ERR_CAPACITY if none is found. */
int ua_puzzle_solve_bounded(ua_ctx *ctx, int n, const uint8_t *tiles, uint64_t mem_bytes, ua_result *res);
/* Blank moves of the last ua_puzzle_solve (0..3 = (+1,0),(-1,0),(0,+1),(0,-1)):
 * replaces GPUPuzzleSolver<N>::getSolution (puzzle/GPU-solver.cuh:276-287). */
int ua_puzzle_moves(ua_ctx *ctx, int32_t *dirs, int64_t cap, int64_t *len);

/* ---- batching: many uneven instances in one persistent launch (no reference counterpart;
 * it generalises the per-query loop of main.cpp:60-110 to many queries at once). */
typedef struct {
    int32_t kind;            /* UA_QUERY_GRID / UA_QUERY_PUZZLE */
    int32_t n;               /* puzzle side */
    const ua_grid *grid;     /* grid queries */
    int32_t sx, sy, ex, ey;
    uint8_t tiles[25];       /* puzzle board, n*n entries */
} ua_query;
int ua_batch_solve(ua_ctx *ctx, const ua_query *qs, int64_t nq, ua_result *out);
/* path / moves of query i of the last ua_batch_solve: walked on demand while its state is
 * resident, and walked before its state is freed when a later sub-batch needs the memory */
int ua_batch_path(ua_ctx *ctx, int64_t i, int32_t *out, int64_t cap, int64_t *len);

/* ---- serving: continuous batching (no reference counterpart). Submitted queries are
 * admitted as team slots and memory free up; ua_serve_step() advances the device by one
 * bounded launch and retires finished queries. Submit/poll are thread-safe. */
int ua_serve_submit(ua_ctx *ctx, const ua_query *q, uint64_t *ticket);
int ua_serve_step(ua_ctx *ctx, int64_t *n_active, int64_t *n_pending);
/* up to cap finished queries: their tickets and results */
int ua_serve_poll(ua_ctx *ctx, uint64_t *tickets, ua_result *results, int64_t cap, int64_t *n_out);
/* Path of a served query (start..goal cell ids, or blank moves 0..3), kept when the
 * context has UA_FLAG_SERVE_PATHS: *len gets its length; with out == NULL only the length
 * is returned, otherwise it is copied out and released. The served counterpart of
 * GPUPathwaySolver::getSolution / GPUPuzzleSolver<N>::getSolution (pathway/GPU-solver.cu:
 * 292-316, puzzle/GPU-solver.cuh:276-287). */
int ua_serve_path(ua_ctx *ctx, uint64_t ticket, int32_t *out, int64_t cap, int64_t *len);
/* steps until nothing is pending or active */
int ua_serve_drain(ua_ctx *ctx);
/* launches so far, their summed duration (HIP events), the device-time budget of one
 * launch in microseconds and the queries widened so far (long-running 15-puzzles move
 * to a team of a quarter of the device) */
int ua_serve_stats(ua_ctx *ctx, int64_t *launches, double *kernel_ms, int32_t *budget_us, int64_t *widened);

/* ---- launch statistics (no reference counterpart: uAstar launches its kernels one by
 * one from a host loop, pathway/GPU-solver.cu:142-290). Every search launch of the
 * context so far -- plain solves, batch sub-batches, serving steps: how many were
 * replayed from a graph and how many issued call by call, their summed host wall time
 * (staging, copies, launch and wait) and their summed kernel time (HIP events). */
int ua_launch_stats(ua_ctx *ctx, int64_t *graph_launches, int64_t *direct_launches, double *wall_ms,
                    double *kernel_ms);
/* why a context created with UA_FLAG_GRAPH stopped replaying launches from graphs
 * ("" while it replays them, and without the flag) */
const char *ua_launch_graph_error(ua_ctx *ctx);

/* ---- autotuning (no reference counterpart: uAstar fixes its launch shape and table
 * sizes at compile time, pathway/GPU-solver.hpp:5-17 and puzzle/GPU-solver.cuh:12-22).
 * ua_autotune runs a caller-given sample of queries under candidate settings and keeps,
 * per instance class, the one with the lowest measured cost: the team width of batched
 * and served grid and 15-puzzle queries, the 15-puzzle hash size (load factor) and the
 * heap count of the plain whole-GPU solve. Nonzero ua_config fields are left alone.
 * The sample is solved several times; the state of an earlier ua_batch_solve is freed. */
typedef struct {
    int32_t grid_team_waves;       /* waves per batched / served grid below grid_large_cells (16, 32, ...) */
    int32_t grid_large_team_waves; /* waves per batched / served grid of at least grid_large_cells cells */
    int64_t grid_large_cells;      /* size threshold between the two grid classes */
    int32_t puzzle_team_waves;     /* waves (= heaps) per batched / served 15-puzzle query */
    int32_t puzzle_hash_log2;      /* hash slots (log2) per batched / served 15-puzzle query */
    int32_t plain_grid_heaps;      /* heaps of the plain whole-GPU grid solve */
    int32_t plain_puzzle_heaps;    /* heaps of the plain whole-GPU 15-puzzle solve */
    int32_t plain_hash_log2;       /* hash slots (log2) of a plain 15-puzzle solve */
    int32_t trials;                /* settings measured by the last ua_autotune */
    double est_grid_ms;            /* estimated stream time of the sample's small grids at the chosen width */
    double est_grid_large_ms;      /* ... of its large grids */
    double est_puzzle_ms;          /* ... of its 15-puzzles */
} ua_tuning;
int ua_autotune(ua_ctx *ctx, const ua_query *sample, int64_t n, ua_tuning *out);
/* the settings in effect (tuned, configured or default) */
int ua_get_tuning(ua_ctx *ctx, ua_tuning *out);
/* applies stored settings without measuring (e.g. from an earlier ua_autotune); a plain
 * heap count of 0 restores the untuned default of its kind */
int ua_set_tuning(ua_ctx *ctx, const ua_tuning *t);

/* ---- diagnostics */
/* Measures random device-memory atomic throughput: blocks x 256 threads each issue iters
 * 32-bit atomicMins (agent scope) at hashed positions of a fresh array of 2^log2_words
 * words; *gops = atomics per nanosecond. The roofline of the search's relaxations. */
int ua_debug_atomic_rate(ua_ctx *ctx, int log2_words, int blocks, int iters, double *gops);
/* Measures the latency of one dependent access: one wave follows a random cycle over
 * 2^log2_words 64-bit words, iters steps of kind 0 (workgroup-scope load), 1 (workgroup-scope
 * atomicMin), 2 (agent-scope atomicMin) or 3 (agent-scope load); *ns = ns per step. */
int ua_debug_access_latency(ua_ctx *ctx, int log2_words, int kind, int iters, double *ns);
/* Runs a random insert/pop sequence on one wave's heap; see tests. */
int ua_debug_heap_selftest(ua_ctx *ctx, const int32_t *ops, int nops, const uint64_t *keys, int64_t nkeys,
                           int cap_nodes, uint64_t *out_pops, int64_t out_cap, int32_t *final_size);

#ifdef __cplusplus
}
#endif
#endif
