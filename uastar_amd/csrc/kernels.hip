// kernels.hip -- the GA* search kernels for gfx950 and their launchers. The team
// search itself (rounds, teams, grid and puzzle expansion) is in search.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "kernels.hpp"
#include "search.hpp"
#include "wave_heap.hpp"

namespace ua {

// The search kernel: a cooperative launch of 16-wave blocks (1024 threads) split into
// groups (see LaunchShape); each group is a team that solves queries of its class
// from a counter. Plain solve: one group of every block and one query. Batch / serve:
// one-block teams for grids (workgroup scope), multi-block groups for 15-puzzles. MODE
// only names the caller (0 plain, 1 batch, 2 serving step), so that profiles report
// the three uses as separate kernels; SET is the problem set (SET_MAIN: grids on bucket
// rings and 3x3 / 4x4 puzzles, SET_FIVE: 5x5 puzzles, SET_GHEAP: grids on per-wave heaps). The shape, the counters, the TeamVars and the
// queries come from the launch's I/O block (StepHeader), so the arguments of a launch
// stay the same from step to step; blocks past the shape's own return at once.
template <int MODE, int SET>
__global__ __launch_bounds__(1024) void k_teams(char *io, uint64_t *segs) {
    StepHeader *h = reinterpret_cast<StepHeader *>(io);
    if (h->shape.iso_xcds == 0 && (int)blockIdx.x >= shape_blocks(h->shape)) return;
    group_block<SET>(reinterpret_cast<const QueryDesc *>(io + h->qd_off),
                     reinterpret_cast<QueryResult *>(io + h->qr_off), h->shape, h->qhead,
                     reinterpret_cast<TeamVars *>(io + h->tv_off), segs);
}

// ------------------------------------------------------------------ path walks
// One lane walks the parent directions from the goal (pathway/GPU-kernel.cuh:546-569
// walks node records instead); ids receive goal..start. Returns the count, or -1 on a
// chain that breaks or runs past cap.
__device__ int64_t walk_grid(const uint64_t *best, int H, int W, int sx, int sy, int ex, int ey, int32_t *ids,
                             int64_t cap) {
    int x = ex, y = ey;
    int64_t c = 0;
    while (c < cap) {
        ids[c++] = x * W + y;
        if (x == sx && y == sy) return c;
        int d = (int)(best[(size_t)x * W + y] & 15ull);
        if (d > 7) break;
        x -= kDX[d];
        y -= kDY[d];
        if (x < 0 || x >= H || y < 0 || y >= W) break;
    }
    return -1;
}

// Blank-move indices goal..start (puzzle/GPU-kernel.cuh:462-509 semantics), for 3x3 / 4x4
// boards of one u64 in 16-byte hash slots (board, complemented value).
template <int N>
__device__ int64_t walk_moves(const uint64_t *slots, int hash_log2, uint64_t goal_slot, int32_t *dirs, int64_t cap) {
    uint64_t hmask = (1ull << hash_log2) - 1;
    uint64_t slot = goal_slot;
    uint64_t b = slots[2 * slot];
    int64_t c = 0;
    while (c <= cap) {
        int d = (int)(~(uint32_t)slots[2 * slot + 1] & 15u);
        if (d == 15) return c;
        if (d > 3 || c == cap) break;
        dirs[c++] = d;
        int bl = nib_pos(b, 0);
        int px = bl / N - kDX[d], py = bl % N - kDY[d];
        if (px < 0 || px >= N || py < 0 || py >= N) break;
        int pb = px * N + py;
        uint64_t t = (b >> (4 * pb)) & 15ull;
        b = b + (t << (4 * bl)) - (t << (4 * pb));
        slot = mix64(b) & hmask;
        int i = 0;
        for (; i < PuzzleProb::MAXPROBE && slots[2 * slot] != b; ++i) slot = (slot + 1) & hmask;
        if (i == PuzzleProb::MAXPROBE) break;
    }
    return -1;
}

// The same walk over two-word 5x5 boards.
__device__ int64_t walk_moves5(const uint64_t *keys, const uint32_t *vals, int hash_log2, uint64_t goal_slot,
                               int32_t *dirs, int64_t cap) {
    uint64_t hmask = (1ull << hash_log2) - 1;
    uint64_t slot = goal_slot;
    uint64_t b0 = keys[2 * slot], b1 = keys[2 * slot + 1];
    int64_t c = 0;
    while (c <= cap) {
        int d = (int)(vals[slot] & 15u);
        if (d == 15) return c;
        if (d > 3 || c == cap) break;
        dirs[c++] = d;
        int bl = pos5(b0, b1, 0);
        int px = bl / 5 - kDX[d], py = bl % 5 - kDY[d];
        if (px < 0 || px >= 5 || py < 0 || py >= 5) break;
        int pb = px * 5 + py;
        uint64_t t = cell5(b0, b1, pb);
        xor5(b0, b1, pb, t);
        xor5(b0, b1, bl, t);
        slot = mix64(b0 ^ mix64(b1)) & hmask;
        int i = 0;
        for (; i < PuzzleProb5::MAXPROBE && (keys[2 * slot] != b0 || keys[2 * slot + 1] != b1); ++i)
            slot = (slot + 1) & hmask;
        if (i == PuzzleProb5::MAXPROBE) break;
    }
    return -1;
}

// Many walks at once, one lane per finished query: job i writes at most jobs[i].len
// entries at out + jobs[i].off and its count (or -1) to jobs[i].got.
__global__ void k_paths(PathJob *jobs, int n, int32_t *out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    PathJob &j = jobs[i];
    int32_t *o = out + j.off;
    if (j.kind == KIND_GRID) j.got = walk_grid(j.best, j.H, j.W, j.sx, j.sy, j.ex, j.ey, o, j.len);
    else if (j.n == 5) j.got = walk_moves5(j.best, j.vals, j.hash_log2, j.goal_slot, o, j.len);
    else if (j.n == 4) j.got = walk_moves<4>(j.best, j.hash_log2, j.goal_slot, o, j.len);
    else j.got = walk_moves<3>(j.best, j.hash_log2, j.goal_slot, o, j.len);
}

// ------------------------------------------------------------------ grid generator
// One thread per bitmap word; see cell_hash() for the spec.
__global__ void k_grid_gen(uint64_t *bm, int H, int W, int wpr, uint64_t seed_mixed, uint64_t thr,
                           int sx, int sy, int ex, int ey) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (size_t)H * wpr) return;
    int x = (int)(w / wpr), y0 = (int)(w % wpr) * 64;
    uint64_t word = 0;
    for (int b = 0; b < 64; ++b) {
        int y = y0 + b;
        if (y >= W) break;
        bool open = (cell_hash(seed_mixed, (uint64_t)x * W + y) >> 32) >= thr;
        if ((x == sx && y == sy) || (x == ex && y == ey)) open = true;
        word |= (uint64_t)open << b;
    }
    bm[w] = word;
}

// Writes the best table of a grid before its search, each word once: all ones for an
// open cell, all ones but the open bit for a blocked one, so that every word carries its
// cell's open bit. Thread t of block (bx, by) takes the cell pair y = 512 bx + 2t, y + 1
// in rows by, by + gridDim.y, ...: both cells lie in one bitmap word (y is even), and a
// pair at an even index of a 16-byte aligned table is one 16-byte store.
__global__ void k_grid_init_best(uint64_t *best, const uint64_t *bm, int H, int W, int wpr, int pair_store) {
    const int y = (int)(blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (y >= W) return;
    for (int x = (int)blockIdx.y; x < H; x += (int)gridDim.y) {
        const uint64_t bits = bm[(size_t)x * wpr + (y >> 6)] >> (y & 63);
        const uint64_t v0 = (bits & 1) ? ~0ull : ~kOpenBit;
        const uint64_t v1 = (bits & 2) ? ~0ull : ~kOpenBit;
        const size_t i = (size_t)x * W + y;
        if (y + 1 >= W) {
            best[i] = v0;
        } else if (pair_store && !(i & 1)) {
            *reinterpret_cast<ulonglong2 *>(best + i) = make_ulonglong2(v0, v1);
        } else {
            best[i] = v0;
            best[i + 1] = v1;
        }
    }
}

// Byte-per-cell input (pathway.hpp:116-119) to the bitmap. Each byte is a cell's mask of
// allowed out-moves (bit d = move d, pathway/CPU-solver.cpp:64-66); the bitmap holds
// all-or-nothing masks only, 0xFF open and 0 blocked, which is what every input module
// writes (input/custom.cpp:23, input/zigzag.cpp:26-39). Any other byte sets *bad, and
// the caller rejects the grid.
__global__ void k_grid_pack(const uint8_t *bytes, uint64_t *bm, int H, int W, int wpr, uint32_t *bad) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (size_t)H * wpr) return;
    int x = (int)(w / wpr), y0 = (int)(w % wpr) * 64;
    uint64_t word = 0;
    bool partial = false;
    for (int b = 0; b < 64; ++b) {
        int y = y0 + b;
        if (y >= W) break;
        const uint8_t v = bytes[(size_t)x * W + y];
        partial |= v != 0 && v != 0xFF;
        word |= (uint64_t)(v != 0) << b;
    }
    bm[w] = word;
    if (partial) atomicOr(bad, 1u);
}

// ------------------------------------------------------------------ PDB build (device BFS)
// Level-synchronous BFS over the k-permutation ranks of one partition, from the goal
// pattern (tile v on cell v-1), a tile moving into any adjacent cell not holding a
// pattern tile at cost 1 (puzzle/database.cpp:37-84). Each thread scans 16 ranks.
template <int N, int P>
__global__ void k_pdb_level(uint8_t *tab, uint64_t size, int d, unsigned long long *added) {
    using S = PSpec<N>;
    constexpr int K = S::K[P];
    constexpr int NN = N * N;
    uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
    unsigned long long cnt = 0;
    if (base < size) {
        uint8_t vals[16];
        if (base + 16 <= size) {
            uint4 v4 = *reinterpret_cast<const uint4 *>(tab + base);
            *reinterpret_cast<uint4 *>(vals) = v4;
        } else {
            for (int i = 0; i < 16; ++i) vals[i] = base + i < size ? tab[base + i] : 0xFF;
        }
        for (int i = 0; i < 16; ++i) {
            if (vals[i] != d) continue;
            uint32_t code = (uint32_t)(base + i);
            int pos[8];
            uint32_t used = 0;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                uint32_t v = code / S::M[P][j];
                code -= v * S::M[P][j];
                int cell = 0;
                for (int c = 0; c < NN; ++c) {  // v-th unused cell
                    if (used >> c & 1u) continue;
                    if (v == 0) {
                        cell = c;
                        break;
                    }
                    --v;
                }
                pos[j] = cell;
                used |= 1u << cell;
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                int x = pos[j] / N, y = pos[j] % N;
#pragma unroll
                for (int dd = 0; dd < 4; ++dd) {
                    int nx = x + kDX[dd], ny = y + kDY[dd];
                    if (nx < 0 || nx >= N || ny < 0 || ny >= N) continue;
                    int np = nx * N + ny;
                    if (used >> np & 1u) continue;
                    uint32_t mask = 0, nc = 0;
#pragma unroll
                    for (int jj = 0; jj < K; ++jj) {
                        int p = jj == j ? np : pos[jj];
                        nc += ((uint32_t)p - __builtin_popcount(mask & ((1u << p) - 1u))) * S::M[P][jj];
                        mask |= 1u << p;
                    }
                    if (tab[nc] == 0xFF) {
                        tab[nc] = (uint8_t)(d + 1);
                        ++cnt;
                    }
                }
            }
        }
    }
    for (int j = 32; j > 0; j >>= 1) cnt += __shfl_xor(cnt, j);
    if (lane_id() == 0 && cnt) atomicAdd(added, cnt);
}

template <int N, int P>
__global__ void k_pdb_seed(uint8_t *tab) {
    using S = PSpec<N>;
    uint32_t mask = 0, code = 0;
    for (int j = 0; j < S::K[P]; ++j) {
        int p = S::T[P][j] - 1;
        code += ((uint32_t)p - __builtin_popcount(mask & ((1u << p) - 1u))) * S::M[P][j];
        mask |= 1u << p;
    }
    tab[code] = 0;
}

// ------------------------------------------------------------------ heap self-test
// Random batches inserted / popped on one wave; pops are written out for the host
// to compare against a sorted reference. ops[i] >= 0: insert keys[k..k+ops[i]);
// ops[i] < 0: pop. out receives 64 keys per pop.
__global__ void k_heap_selftest(const int *ops, int nops, const uint64_t *keys, uint64_t *storage,
                                int cap, uint64_t *out, int *status) {
    WaveHeap h;
    h.init(storage, cap);
    int l = lane_id();
    int k = 0, o = 0;
    bool ok = true;
    for (int i = 0; i < nops; ++i) {
        int c = ops[i];
        if (c >= 0) {
            uint64_t v = l < c ? keys[k + l] : SENT;
            k += c;
            v = wave_sort(v);
            ok &= h.insert(v, c);
        } else {
            out[(size_t)o * 64 + l] = h.pop();
            ++o;
        }
    }
    if (l == 0) *status = ok ? h.size() : -1;
}

// Throughput probe of random device-memory atomics: every thread issues `iters`
// atomicMins at agent scope on 32-bit words of a[0 .. mask] picked by a hash of (thread,
// i), the access pattern of the search's relaxations. The sum of the old values goes to
// *sink only if it hits a magic value, so that no result is dead.
__global__ __launch_bounds__(256) void k_atomic_probe(uint32_t *a, uint64_t mask, int iters, uint32_t *sink) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (int i = 0; i < iters; ++i) {
        const uint64_t idx = mix64(t * 0x9e3779b97f4a7c15ull + (uint64_t)i) & mask;
        acc += __hip_atomic_fetch_min(a + idx, (uint32_t)(t + (uint64_t)i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (acc == 0x12345678u) sink[0] = acc;
}

// Latency probe: one wave follows a random cycle of indices (a[i] holds the next index),
// one dependent access a step: KIND 0 a workgroup-scope load, 1 a workgroup-scope
// atomicMin that changes nothing, 2 the same at agent scope, 3 an agent-scope load. The
// atomicMin's operand, mask + 1, exceeds every entry but is a run-time value: a constant
// all-ones operand would let the compiler turn the atomic into a load.
// out[0] = s_memrealtime ticks (100 MHz) of `iters` steps, out[1] the last index.
template <int KIND>
__global__ __launch_bounds__(64) void k_latency_probe(uint64_t *a, uint64_t mask, int iters, uint64_t *out) {
    uint64_t idx = ((uint64_t)threadIdx.x * 977u) & mask;
    const int64_t t0 = (int64_t)__builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < iters; ++i) {
        uint64_t v;
        if (KIND == 0) v = __hip_atomic_load(a + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (KIND == 1) v = __hip_atomic_fetch_min(a + idx, mask + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (KIND == 2) v = __hip_atomic_fetch_min(a + idx, mask + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else v = __hip_atomic_load(a + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        idx = v & mask;
    }
    const int64_t t1 = (int64_t)__builtin_amdgcn_s_memrealtime();
    if (threadIdx.x == 0) {
        out[0] = (uint64_t)(t1 - t0);
        out[1] = idx;
    }
}

// ------------------------------------------------------------------ launchers
static int grid_blocks(size_t n, int t) { return (int)((n + t - 1) / t); }

hipError_t launch_latency_probe(uint64_t *a, uint64_t mask, int kind, int iters, uint64_t *out, hipStream_t st) {
    if (kind == 0) hipLaunchKernelGGL(k_latency_probe<0>, dim3(1), dim3(64), 0, st, a, mask, iters, out);
    else if (kind == 1) hipLaunchKernelGGL(k_latency_probe<1>, dim3(1), dim3(64), 0, st, a, mask, iters, out);
    else if (kind == 2) hipLaunchKernelGGL(k_latency_probe<2>, dim3(1), dim3(64), 0, st, a, mask, iters, out);
    else hipLaunchKernelGGL(k_latency_probe<3>, dim3(1), dim3(64), 0, st, a, mask, iters, out);
    return hipGetLastError();
}

hipError_t launch_atomic_probe(uint32_t *a, uint64_t mask, int blocks, int iters, uint32_t *sink, hipStream_t st) {
    hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, st, a, mask, iters, sink);
    return hipGetLastError();
}

hipError_t launch_grid_gen(uint64_t *bm, int H, int W, int wpr, uint64_t seed, uint64_t thr, int sx,
                           int sy, int ex, int ey, hipStream_t st) {
    size_t n = (size_t)H * wpr;
    hipLaunchKernelGGL(k_grid_gen, dim3(grid_blocks(n, 256)), dim3(256), 0, st, bm, H, W, wpr,
                       mix64(seed), thr, sx, sy, ex, ey);
    return hipGetLastError();
}

hipError_t launch_grid_pack(const uint8_t *bytes, uint64_t *bm, int H, int W, int wpr, uint32_t *bad,
                            hipStream_t st) {
    size_t n = (size_t)H * wpr;
    hipLaunchKernelGGL(k_grid_pack, dim3(grid_blocks(n, 256)), dim3(256), 0, st, bytes, bm, H, W, wpr, bad);
    return hipGetLastError();
}

hipError_t launch_grid_init_best(uint64_t *best, const uint64_t *bm, int H, int W, int wpr, hipStream_t st) {
    if (H <= 0 || W <= 0) return hipSuccess;
    const dim3 grid((unsigned)((W + 511) / 512), (unsigned)std::min(H, 65535));
    const int pair_store = ((uintptr_t)best & 15) == 0;
    hipLaunchKernelGGL(k_grid_init_best, grid, dim3(256), 0, st, best, bm, H, W, wpr, pair_store);
    return hipGetLastError();
}

hipError_t teams_max_blocks(int *per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_teams<0, SET_MAIN>, kBlockWaves * 64, 0);
}

const void *teams_fn(int mode, int set) {
    if (set == SET_FIVE)
        return mode == 2 ? (const void *)k_teams<2, SET_FIVE> : mode == 1 ? (const void *)k_teams<1, SET_FIVE>
                                                                          : (const void *)k_teams<0, SET_FIVE>;
    if (set == SET_GHEAP)
        return mode == 2 ? (const void *)k_teams<2, SET_GHEAP> : mode == 1 ? (const void *)k_teams<1, SET_GHEAP>
                                                                           : (const void *)k_teams<0, SET_GHEAP>;
    return mode == 2 ? (const void *)k_teams<2, SET_MAIN> : mode == 1 ? (const void *)k_teams<1, SET_MAIN>
                                                                      : (const void *)k_teams<0, SET_MAIN>;
}

hipError_t launch_teams(char *io, int blocks, uint64_t *segs, int mode, int set, hipStream_t st) {
    void *args[] = {(void *)&io, (void *)&segs};
    return hipLaunchCooperativeKernel(teams_fn(mode, set), dim3(blocks), dim3(kBlockWaves * 64), args, 0, st);
}

hipError_t launch_paths(PathJob *jobs, int n, int32_t *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_paths, dim3(grid_blocks((size_t)n, 64)), dim3(64), 0, st, jobs, n, out);
    return hipGetLastError();
}

template <int N, int P>
static hipError_t pdb_one(uint8_t *tab, uint64_t size, unsigned long long *d_added, hipStream_t st) {
    hipError_t e = hipMemsetAsync(tab, 0xFF, size, st);
    if (e) return e;
    hipLaunchKernelGGL((k_pdb_seed<N, P>), dim3(1), dim3(1), 0, st, tab);
    for (int d = 0; d < 250; ++d) {
        e = hipMemsetAsync(d_added, 0, sizeof(*d_added), st);
        if (e) return e;
        uint64_t threads = (size + 15) / 16;
        hipLaunchKernelGGL((k_pdb_level<N, P>), dim3(grid_blocks(threads, 256)), dim3(256), 0, st, tab,
                           size, d, d_added);
        unsigned long long h = 0;
        e = hipMemcpyAsync(&h, d_added, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e) return e;
        e = hipStreamSynchronize(st);
        if (e) return e;
        if (!h) break;
    }
    return hipGetLastError();
}

hipError_t build_pdb(int n, uint8_t *pdb, unsigned long long *d_added, hipStream_t st) {
    if (n == 5) {
        const uint64_t sz = 127512000ull;  // 25!/19!
        hipError_t e = pdb_one<5, 0>(pdb + PSpec<5>::OFF[0], sz, d_added, st);
        if (!e) e = pdb_one<5, 1>(pdb + PSpec<5>::OFF[1], sz, d_added, st);
        if (!e) e = pdb_one<5, 2>(pdb + PSpec<5>::OFF[2], sz, d_added, st);
        if (!e) e = pdb_one<5, 3>(pdb + PSpec<5>::OFF[3], sz, d_added, st);
        return e;
    }
    if (n == 4) {
        hipError_t e = pdb_one<4, 0>(pdb, PSpec<4>::OFF[1], d_added, st);
        if (e) return e;
        return pdb_one<4, 1>(pdb + PSpec<4>::OFF[1], 57657600ull, d_added, st);
    }
    return pdb_one<3, 0>(pdb, 362880ull, d_added, st);
}

hipError_t launch_heap_selftest(const int *ops, int nops, const uint64_t *keys, uint64_t *storage, int cap,
                                uint64_t *out, int *status, hipStream_t st) {
    hipLaunchKernelGGL(k_heap_selftest, dim3(1), dim3(64), 0, st, ops, nops, keys, storage, cap, out, status);
    return hipGetLastError();
}

}  // namespace ua
