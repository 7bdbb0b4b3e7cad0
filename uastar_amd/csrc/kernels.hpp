// kernels.hpp -- host-callable launchers of kernels.hip (internal to libuastar).
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace ua {

hipError_t launch_grid_gen(uint64_t *bm, int H, int W, int wpr, uint64_t seed, uint64_t thr, int sx,
                           int sy, int ex, int ey, hipStream_t st);
// Bytes (0 blocked, 0xFF open) to the bitmap; *bad (zeroed by the caller) becomes nonzero
// if any byte is a partial out-move mask.
hipError_t launch_grid_pack(const uint8_t *bytes, uint64_t *bm, int H, int W, int wpr, uint32_t *bad,
                            hipStream_t st);
// Writes a grid's best table (H * W words) before its search: all ones for an open cell,
// all ones but the open bit for a blocked one.
hipError_t launch_grid_init_best(uint64_t *best, const uint64_t *bm, int H, int W, int wpr, hipStream_t st);
// The search kernel (k_teams): `blocks` blocks of kBlockWaves wavefronts, cooperative,
// over the I/O block io (StepHeader, then TeamVars, descriptors and results). mode
// names the caller in profiles: 0 plain solve, 1 batch, 2 serving step; set is the
// problem set (0: grids and 3x3 / 4x4 puzzles, 1: 5x5 puzzles). teams_fn is the kernel
// of a mode and set, for a graph's kernel node (arguments io, segs).
hipError_t teams_max_blocks(int *per_cu);
const void *teams_fn(int mode, int set);
hipError_t launch_teams(char *io, int blocks, uint64_t *segs, int mode, int set, hipStream_t st);
// Path walks of finished queries, one lane each (see PathJob).
hipError_t launch_paths(PathJob *jobs, int n, int32_t *out, hipStream_t st);
hipError_t build_pdb(int n, uint8_t *pdb, unsigned long long *d_added, hipStream_t st);
hipError_t launch_latency_probe(uint64_t *a, uint64_t mask, int kind, int iters, uint64_t *out, hipStream_t st);
hipError_t launch_atomic_probe(uint32_t *a, uint64_t mask, int blocks, int iters, uint32_t *sink, hipStream_t st);
hipError_t launch_heap_selftest(const int *ops, int nops, const uint64_t *keys, uint64_t *storage,
                                int cap, uint64_t *out, int *status, hipStream_t st);

}  // namespace ua
