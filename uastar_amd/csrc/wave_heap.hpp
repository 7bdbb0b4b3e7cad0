// wave_heap.hpp -- wave64 primitives and the per-wavefront batched binary heap.
//
// The open list of each heap is a binary heap whose nodes are sorted batches of
// 64 keys (one key per lane, coalesced 512-byte loads and stores). The top of the
// heap -- the root batch and a partial batch of fewer than 64 keys -- stays in
// registers of the owning wavefront; nodes 2..n live in HBM. Invariants:
//   (H1) for every node v and child c: max(v) <= min(c)
//   (H2) every key of the partial batch is >= max(root)
// so the root holds the 64 smallest keys of the heap and a pop returns them.
// All merges are wave-wide bitonic networks over registers (DPP and permlane swaps).
// This replaces uAstar's one-heap-per-thread sift-down/sift-up in global memory
// (pathway/GPU-kernel.cuh:209-252 and :517-535; puzzle/GPU-kernel.cuh:262-298, :435-453).
#pragma once
#ifndef UA_EMU
#include <hip/hip_runtime.h>
#endif
#include "common.hpp"

namespace ua {

// Pointers into device memory (HBM) as the global address space: their loads, stores and
// atomics compile to global_* instructions (SGPR base plus a lane offset where the base is
// uniform) instead of flat ones, which also count against the LDS counter, so that a wait
// for an LDS result does not wait for HBM loads still in flight.
#ifdef UA_EMU
#define UA_GLOBAL
#define UA_LDS
#else
#define UA_GLOBAL __attribute__((address_space(1)))
#define UA_LDS __attribute__((address_space(3)))
#endif
template <class T>
__device__ __forceinline__ UA_GLOBAL T *gptr(T *p) {
    return (UA_GLOBAL T *)p;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    int lo = __shfl((int)(uint32_t)v, src);
    int hi = __shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
#ifdef UA_EMU
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    int lo = __shfl_xor((int)(uint32_t)v, m);
    int hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
#else
// Lane exchange l <-> l ^ m of one dword without the LDS crossbar (m a compile-time
// constant after unrolling): DPP within rows of 16 lanes, the gfx950 permlane swaps
// across rows and halves.
//   m = 1, 2: quad_perm [1,0,3,2] / [2,3,0,1]
//   m = 4:    row_half_mirror (l -> 7 - l within 8), then quad_perm [3,2,1,0]
//   m = 8:    row_mirror (l -> 15 - l within 16), then row_half_mirror
//   m = 16:   permlane16_swap of v with itself: the first result holds rows (0,0,2,2),
//             the second rows (1,1,3,3)
//   m = 32:   permlane32_swap of v with itself: the first holds halves (0,0), the second (1,1)
__device__ __forceinline__ uint32_t xor_lanes32(uint32_t v, int m) {
    const int x = (int)v;
    switch (m) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    case 4: return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false), 0x1B, 0xF,
                                                      0xF, false);
    case 8: return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false), 0x141, 0xF,
                                                      0xF, false);
    case 16: {
        auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (__lane_id() & 16) ? p[0] : p[1];
    }
    case 32: {
        auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (__lane_id() & 32) ? p[0] : p[1];
    }
    default: return (uint32_t)__shfl_xor(x, m);
    }
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = xor_lanes32((uint32_t)v, m), hi = xor_lanes32((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
#endif
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
    uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a < b ? b : a; }

// One compare-exchange between lanes l and l ^ J (J a power of two below 64): a lane
// whose want_min is set keeps the smaller key of the pair, the other lane the larger.
// One 64-bit compare and two selects per stage. For J = 16 and 32 the gfx950 permlane
// swaps of the key with itself hand every lane both keys of its pair (x from the pair's
// lower lane, y from its upper lane), so no lane-dependent pick of the partner is needed.
template <int J>
__device__ __forceinline__ uint64_t cmpx(uint64_t v, bool want_min) {
#ifndef UA_EMU
    if constexpr (J == 16 || J == 32) {
        const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
        auto pl = J == 16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                          : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        auto ph = J == 16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                          : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        const uint64_t x = ((uint64_t)ph[0] << 32) | pl[0];
        const uint64_t y = ((uint64_t)ph[1] << 32) | pl[1];
        return ((y < x) == want_min) ? y : x;
    }
#endif
    const uint64_t o = shfl_xor64(v, J);
    return ((o < v) == want_min) ? o : v;
}

// Ascending bitonic sort of one key per lane.
__device__ __forceinline__ uint64_t wave_sort(uint64_t v) {
    const int l = lane_id();
#define UA_SORT_STAGE(K, J) v = cmpx<J>(v, ((l & (J)) == 0) == ((l & (K)) == 0))
    UA_SORT_STAGE(2, 1);
    UA_SORT_STAGE(4, 2); UA_SORT_STAGE(4, 1);
    UA_SORT_STAGE(8, 4); UA_SORT_STAGE(8, 2); UA_SORT_STAGE(8, 1);
    UA_SORT_STAGE(16, 8); UA_SORT_STAGE(16, 4); UA_SORT_STAGE(16, 2); UA_SORT_STAGE(16, 1);
    UA_SORT_STAGE(32, 16); UA_SORT_STAGE(32, 8); UA_SORT_STAGE(32, 4); UA_SORT_STAGE(32, 2); UA_SORT_STAGE(32, 1);
    UA_SORT_STAGE(64, 32); UA_SORT_STAGE(64, 16); UA_SORT_STAGE(64, 8); UA_SORT_STAGE(64, 4); UA_SORT_STAGE(64, 2);
    UA_SORT_STAGE(64, 1);
#undef UA_SORT_STAGE
    return v;
}

// In: a, b ascending. Out: a = the 64 smallest of a U b, b = the 64 largest, both ascending.
// Lane l first compares a[l] with b[63 - l]; the two halves are then bitonic and each
// is cleaned by six half-cleaner stages.
__device__ __forceinline__ void wave_merge_net(uint64_t &a, uint64_t &b) {
    const int l = lane_id();
    const uint64_t br = shfl64(b, 63 - l);
    const bool c = br < a;
    uint64_t lo = c ? br : a, hi = c ? a : br;
#define UA_MERGE_STAGE(J)                 \
    do {                                  \
        const bool wm = (l & (J)) == 0;   \
        lo = cmpx<J>(lo, wm);             \
        hi = cmpx<J>(hi, wm);             \
    } while (0)
    UA_MERGE_STAGE(32);
    UA_MERGE_STAGE(16);
    UA_MERGE_STAGE(8);
    UA_MERGE_STAGE(4);
    UA_MERGE_STAGE(2);
    UA_MERGE_STAGE(1);
#undef UA_MERGE_STAGE
    a = lo;
    b = hi;
}
// The same, skipping the network when the two batches do not overlap (wave-uniform
// test on two lanes): a <= b already, or b <= a and the batches only trade places.
__device__ __forceinline__ void wave_merge(uint64_t &a, uint64_t &b) {
    if (readlane64(a, 63) <= readlane64(b, 0)) return;
    if (readlane64(b, 63) <= readlane64(a, 0)) {
        const uint64_t t = a;
        a = b;
        b = t;
        return;
    }
    wave_merge_net(a, b);
}

// Exclusive prefix sum of a per-lane count; returns the wave total in *total.
__device__ __forceinline__ int wave_excl_scan(int c, int *total) {
    const int l = lane_id();
    int s = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(s, d);
        if (l >= d) s += o;
    }
    *total = __builtin_amdgcn_readlane(s, 63);
    return s - c;
}

__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) v = umin64(v, shfl_xor64(v, j));
    return readlane64(v, 0);
}
__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) v += (int64_t)shfl_xor64((uint64_t)v, j);
    return (int64_t)readlane64((uint64_t)v, 0);
}

// ------------------------------------------------------------------ the heap
struct WaveHeap {
    uint64_t root;    // this lane's key of the root batch
    uint64_t part;    // this lane's key of the partial batch (SENT beyond pc)
    int n;            // full nodes, root included (wave-uniform)
    int pc;           // keys in the partial batch (wave-uniform)
    int cap;          // max full nodes, root included
    uint64_t *nodes;  // node i >= 2 at nodes + (i - 2) * 64

    __device__ void init(uint64_t *storage, int capacity) {
        root = SENT;
        part = SENT;
        n = 0;
        pc = 0;
        cap = capacity;
        nodes = storage;
    }
    __device__ __forceinline__ uint64_t load(int i) const {
        return gptr(nodes)[(size_t)(i - 2) * 64 + lane_id()];
    }
    __device__ __forceinline__ void store(int i, uint64_t v) const {
        gptr(nodes)[(size_t)(i - 2) * 64 + lane_id()] = v;
    }
    __device__ __forceinline__ uint64_t min_key() const {
        return n >= 1 ? readlane64(root, 0) : readlane64(part, 0);
    }
    __device__ __forceinline__ int size() const { return n * 64 + pc; }

    // Pops the (up to) 64 smallest keys, one per lane, ascending; SENT where empty.
    //
    // The root becomes a hole that descends: at each level the hole's two children
    // merge, the low half fills the hole and the high half replaces the child with the
    // larger max (its subtree's keys are at least that max); the other child is the new
    // hole. One merge a level, with the next level's children loaded before it. At a
    // leaf the hole takes the last node, which moves up while its parent's max exceeds
    // its min (rarely: the last node holds late, large keys). Finally the partial batch
    // gives up any keys below the new root's max (H2).
    __device__ uint64_t pop() {
        uint64_t out;
        if (n >= 1) {
            out = root;
            if (n >= 2) {
                int v = 1;
                uint64_t L = load(2);
                uint64_t R = n >= 3 ? load(3) : SENT;
                while (true) {
                    const int l = 2 * v;  // L, R hold nodes l and l + 1 (R = SENT beyond n)
                    if (l + 1 > n) {      // a single child: it moves up into the hole
                        if (v == 1) root = L;
                        else store(v, L);
                        v = l;
                        break;
                    }
                    const uint64_t lmax = readlane64(L, 63), rmax = readlane64(R, 63);
                    const int big = lmax >= rmax ? l : l + 1;
                    const int small = big == l ? l + 1 : l;
                    const int nl = 2 * small;
                    const uint64_t NL = nl <= n ? load(nl) : SENT;
                    const uint64_t NR = nl + 1 <= n ? load(nl + 1) : SENT;
                    wave_merge(L, R);  // L = low half, R = high half
                    store(big, R);
                    if (v == 1) root = L;
                    else store(v, L);
                    v = small;
                    if (nl > n) break;  // the hole is a leaf
                    L = NL;
                    R = NR;
                }
                if (v != n) {  // fill the hole with the last node and let it rise
                    uint64_t Y = load(n);
                    --n;
                    int u = v;
                    while (u > 1) {
                        const int p = u >> 1;
                        uint64_t P = p == 1 ? root : load(p);
                        if (readlane64(P, 63) <= readlane64(Y, 0)) break;
                        wave_merge(P, Y);  // P = low half (rises), Y = high half (stays at u)
                        store(u, Y);
                        u = p;
                        Y = P;
                    }
                    if (u == 1) root = Y;
                    else store(u, Y);
                } else {
                    --n;
                }
                wave_merge(root, part);  // root keeps the 64 smallest, partial keeps pc keys
            } else {
                n = 0;
                root = SENT;
            }
        } else {
            out = part;
            part = SENT;
            pc = 0;
        }
        return out;
    }

    // Inserts a full batch F whose keys are all >= max(root), as a new leaf,
    // merging top-down along the path (the root is skipped). The path's nodes are
    // loaded PF at a time before any merge, so a deep heap costs few load latencies.
    // Returns false on overflow.
    __device__ bool push_full(uint64_t F) {
        constexpr int PF = 4;
        int target = n + 1;
        if (target > cap) return false;
        int depth = 31 - __builtin_clz((unsigned)target);
        for (int L0 = 1; L0 < depth; L0 += PF) {
            uint64_t X[PF];
#pragma unroll
            for (int j = 0; j < PF; ++j)
                if (L0 + j < depth) X[j] = load(target >> (depth - L0 - j));
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                if (L0 + j >= depth) break;
                if (readlane64(X[j], 63) <= readlane64(F, 0)) continue;
                wave_merge(X[j], F);
                store(target >> (depth - L0 - j), X[j]);
            }
        }
        store(target, F);
        n = target;
        return true;
    }

    // Inserts s (<= 64) keys S, ascending with SENT padding. Returns false on overflow.
    __device__ bool insert(uint64_t S, int s) {
        if (s <= 0) return true;
        int t = pc + s;
        if (n >= 1) {
            wave_merge(root, S);        // root keeps the 64 smallest, S the s largest
            wave_merge(part, S);        // part = smallest min(64,t), S = the rest
            if (t < 64) {
                pc = t;
            } else {
                uint64_t F = part;      // 64 keys, all >= max(root)
                part = S;
                pc = t - 64;
                return push_full(F);
            }
        } else {
            wave_merge(part, S);
            if (t < 64) {
                pc = t;
            } else {
                root = part;
                part = S;
                pc = t - 64;
                n = 1;
            }
        }
        return true;
    }
};

}  // namespace ua
