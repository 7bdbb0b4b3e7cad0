/*
 * oracle.h -- CPU restatement of uAstar's CPU solvers.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (uastar_amd/, include/,
 * the CLI) may include, link or call this. Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only
 * as the checker / the timed CPU baseline.
 *
 * What each function restates (paths relative to the uAstar source tree):
 *   orc_grid_astar  variant 0/2  pathway/CPU-solver.cpp:46-112 (A*, lazy deletion,
 *                                source-cell mask rule, strict-< relax, goal on pop),
 *                                utils.hpp:119-122 (move order and costs)
 *   orc_pdb_*                    puzzle/database.cpp:8-199 (BFS, k-permutation rank)
 *   orc_puzzle_astar             puzzle/CPU-solver.hpp:43-314 (partitions, A*, moves)
 * Everything else (generators, Dijkstra, IDA*, dense A*) is a cross-check or a
 * project-defined input spec; see the comments at each definition.
 */
#ifndef UASTAR_ORACLE_H
#define UASTAR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t found;        /* 1 = path found */
    int64_t orth;         /* exact cost = orth + diag*sqrt(2) (variants 0,1,3) */
    int64_t diag;
    float cost_f32;       /* variant 2: the reference's FP32 cost; others: orth + diag*SQRT2f */
    int64_t expanded;     /* nodes closed (popped and expanded, goal included) */
    int64_t generated;    /* legal moves relaxed */
    int64_t inserted;     /* open-list pushes (start included) */
    int64_t path_len;     /* number of cells on the path, start and goal included */
} orc_grid_out;

/* ---- project-defined input generators (same spec as the product's, coded independently) ---- */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_cell_hash(uint64_t seed, uint64_t id);
/* bytes[x*W+y] = 0xFF (open) or 0 (blocked); blocked iff (cell_hash>>32) < thr; start/goal forced open */
void orc_grid_bernoulli(int H, int W, uint64_t seed, uint64_t thr,
                        int sx, int sy, int ex, int ey, uint8_t *bytes);
/* pathway/input/zigzag.cpp:14-40 */
void orc_grid_zigzag(int H, int W, int gap, uint8_t *bytes, int *sx, int *sy, int *ex, int *ey);
/* random solvable N*N board, row-major, 0 = blank */
void orc_puzzle_random(int n, uint64_t seed, uint8_t *tiles);
int orc_puzzle_solvable(int n, const uint8_t *tiles);

/* ---- grid search ----
 * graph: one byte per cell, bit i = move i allowed out of the cell (pathway.hpp:116-119).
 * variant 0: exact (orth,diag) costs with the reference's containers
 *            (priority_queue + unordered_set/unordered_map)
 * variant 1: exact costs, dense arrays (for grids too large for hash maps)
 * variant 2: the reference's FP32 costs, reference containers
 * variant 3: Dijkstra (h = 0), exact costs, dense arrays (textbook cross-check)
 * path: receives cell ids start..goal if path != NULL and the path fits path_cap.
 * Returns 0 on success, <0 on bad arguments. */
int orc_grid_astar(const uint8_t *graph, int H, int W, int sx, int sy, int ex, int ey,
                   int variant, orc_grid_out *out, int32_t *path, int64_t path_cap);

/* ---- pattern databases (puzzle/database.cpp) ---- */
uint64_t orc_pdb_size(int n, int k);
uint32_t orc_pdb_encode(int n, const int *tracked, int k, const uint8_t *conf);
void orc_pdb_decode(int n, const int *tracked, int k, uint32_t code, uint8_t *conf);
/* method 0: the reference's queue BFS (database.cpp:37-84), single thread
 * method 1: level-synchronous BFS over OpenMP threads (same table)            */
int orc_pdb_build(int n, const int *tracked, int k, int method, uint8_t *out);
/* reference partitions (puzzle/CPU-solver.hpp:44-104): count, sizes, tiles */
int orc_partitions(int n, int *counts, int *tiles /* [4][8] */);

/* ---- puzzle search ----
 * pdb: concatenation of the partition tables in partition order (GPU-solver.cuh:459-467).
 * variant 0: A* with the reference's containers (CPU-solver.hpp:154-231)
 * variant 1: IDA* with the same PDB heuristic (cross-check)
 * variant 2: IDA* with Manhattan distance (textbook cross-check, easy boards only)
 * moves: blank move indices 0..3 = (+1,0),(-1,0),(0,+1),(0,-1) (utils.hpp:120-121).
 * Returns optimal steps (>=0), -1 if no solution, -2 on bad args, -3 node limit. */
int orc_puzzle_astar(int n, const uint8_t *tiles, const uint8_t *pdb, int variant,
                     int64_t node_limit, int64_t *expanded, int32_t *moves, int64_t moves_cap,
                     int64_t *moves_len);
/* h(board) as the reference's computeHValue (CPU-solver.hpp:287-314) */
int orc_puzzle_h(int n, const uint8_t *tiles, const uint8_t *pdb);

#ifdef __cplusplus
}
#endif
#endif
