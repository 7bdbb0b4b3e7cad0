// search.hpp -- the GA* team search (device code shared by every search kernel).
//
// One search instance is run by a *team* of wavefronts. Each wavefront owns one
// batched binary heap (wave_heap.hpp). A round is
//   phase A  every wave pops its 64 smallest keys, expands them, relaxes the
//            successors with one 64-bit atomicMin each (grid: dense best-cost
//            table; puzzle: lock-free open-addressing hash), and appends the
//            improved successors to the team's insert buffer. On grids a lane whose
//            improvements are one equal-f child and at most one other expands that
//            child in the same round (GridProb::pick_child);
//   barrier  (the last arriver resets the next round's slot)
//   phase B  every wave reads the round's minimum popped key (LB) and best goal;
//            the team stops when goal <= LB; otherwise each wave takes a strided
//            share of the insert buffer, sorts it and inserts it into its heap.
// This restates the round of uAstar's kExtractExpand / kAssign / kDeduplicate /
// kHeapInsert (pathway/GPU-kernel.cuh:169-544; puzzle/GPU-kernel.cuh:215-460) and the
// host loop's termination test (puzzle/GPU-solver.cuh:243-251), but with the test
// on the device every round and with duplicate detection done by atomics instead
// of a sort: an improvement is pushed exactly when the atomicMin lowered the cost.
//
// Teams: a GroupTeam is a group of blocks (16 waves each) of one cooperative launch.
// The plain one-query-per-launch solve uses one group of plain_heaps / 16 blocks;
// batched and served queries use a team of their class's width (one block by default
// for grids, eight for 15-puzzles) per query, pulled from a work counter.
#pragma once
#ifndef UA_EMU
#include <hip/hip_runtime.h>
#endif

#include "common.hpp"
#include "wave_heap.hpp"

namespace ua {

// ------------------------------------------------------------------ teams
// A team owns: the two round slots, the two halves of the insert buffer (segcap
// keys each) and four statistics counters.
struct TeamMem {
    RoundSlot *slot;   // [2]
    RoundSlot *view;   // LDS: the round slot as read by one lane, for the whole block
    uint64_t *seg;     // [2][segcap]
    int segcap;
    uint64_t *stats;   // [5]: pops, expanded, generated, inserted, atomics
    void *aux;         // LDS of a one-block team (bucket counters of a bucket open list), else null
};

// Relaxed atomics at a scope chosen at run time. A one-block team is the only user of
// its query's data -- best-cost table or hash, heaps, round slots, insert buffer --
// for the whole launch, so it exchanges all of it at workgroup scope (atomics resolve
// in this XCD's L2; barriers need no L2 write-back or invalidation). A team of several
// blocks uses agent scope. Between launches the kernel boundary makes the state visible
// wherever a resumed query lands. The flag is wave-uniform: each operation is one
// scalar branch, and one copy of the search serves both kinds of team (two inlined
// copies would double the kernel's register use and halve its occupancy).
struct Scope {
    bool wg;
    template <class T>
    __device__ __forceinline__ T load(const T *p) const {
        return wg ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // A read that may return an older value (workgroup scope: served by this CU's
    // caches, which nothing refreshes between the rounds of a team of several blocks).
    // For the monotone search state only -- best words and hash values only decrease, a
    // hash key changes once from 0 -- where an older value costs at most a failing or
    // needless atomic, never a wrong answer. The state of a popped key is read with
    // load(), whose value is at least as new as the one that pushed the key.
    template <class T>
    __device__ __forceinline__ T load_weak(const T *p) const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <class T>
    __device__ __forceinline__ void store(T *p, T v) const {
        if (wg) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T fmin(T *p, T v) const {
        return wg ? __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T fmax(T *p, T v) const {
        return wg ? __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T fadd(T *p, T v) const {
        return wg ? __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ void fold_or(T *p, T v) const {
        if (wg) __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // compare-and-swap; on failure `expect` receives the current value
    __device__ __forceinline__ bool cas(uint64_t *p, uint64_t &expect, uint64_t v) const {
        return wg ? __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
    }
#ifndef UA_EMU
    // The same operations on pointers known to be in device memory (global instructions).
    template <class T>
    __device__ __forceinline__ T load(const UA_GLOBAL T *p) const {
        return wg ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T load_weak(const UA_GLOBAL T *p) const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <class T>
    __device__ __forceinline__ void store(UA_GLOBAL T *p, T v) const {
        if (wg) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T fmin(UA_GLOBAL T *p, T v) const {
        return wg ? __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <class T>
    __device__ __forceinline__ T fmax(UA_GLOBAL T *p, T v) const {
        return wg ? __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ bool cas(UA_GLOBAL uint64_t *p, uint64_t &expect, uint64_t v) const {
        return wg ? __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP)
                  : __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
    }
#endif
    __device__ __forceinline__ void release() const {
        if (wg) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
    __device__ __forceinline__ void acquire() const {
        if (wg) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
};

__device__ __forceinline__ int64_t now_ticks() { return (int64_t)__builtin_amdgcn_s_memrealtime(); }

// Per-lane search counters of one launch (32-bit: a lane's counts stay below 2^32 for any
// round limit), summed over the team at the end of a query:
//   exp  states expanded (pops that were not stale)
//   gen  legal moves generated
//   ins  successors whose cost the lane's atomic lowered (inserted or improved)
//   atom device-memory atomics issued on the search state (relaxing min, hash-claiming CAS)
struct Ctr {
    uint32_t exp = 0, gen = 0, ins = 0, atom = 0;
};

// A group of `bpg` consecutive blocks of a cooperative launch. Round state and the
// insert buffer live in global memory and cross a generation barrier over the
// group's blocks. The plain solve is one group of every block; a batch is many groups
// (one block each for grids), each pulling queries from a shared counter.
struct GroupTeam {
    static constexpr uint32_t SPIN_LIMIT = 1u << 22;  // a few seconds; then the group aborts
    int tw, W;      // wave index within the group (wave-uniform), waves in the group
    int gblock;     // block index within the group
    int bpg;        // blocks per group
    Scope sc;       // workgroup scope for one-block groups, agent scope otherwise
    TeamMem m;
    TeamVars *tv;
    __device__ GroupTeam(const TeamMem &mem, TeamVars *t, int blocks_per_group, int block_in_group)
        : m(mem), tv(t) {
        bpg = blocks_per_group;
        gblock = block_in_group;
        sc.wg = bpg == 1;
        int wpb = (int)(blockDim.x >> 6);
        tw = __builtin_amdgcn_readfirstlane(gblock * wpb + (int)(threadIdx.x >> 6));
        W = bpg * wpb;
    }
    __device__ bool leader() const { return gblock == 0 && threadIdx.x == 0; }
    // Generation barrier. Every block's thread 0 arrives once; the last one runs the
    // hook, resets the counter and bumps the generation. A bounded spin turns a lost
    // block into an abort that every block of the group observes, never a hang.
    // A one-block group is one workgroup barrier around the hook.
    //
    // Across blocks, an agent-scope fence would cost an L2 write-back (release) or an
    // L1/L2 invalidation (acquire) on this XCD, which also evicts the data of the other
    // blocks there. So a group of several blocks exchanges everything its blocks share
    // -- insert buffer, round slots, hash, best words, counters -- only through
    // agent-scope atomics (relaxed loads and stores at agent scope go to the device's
    // coherence point, past this XCD's L2), and the barrier orders them without any
    // agent-scope fence: every wave waits until its own memory operations are complete
    // (s_waitcnt 0) before the workgroup barrier and thread 0's arrival, and every read
    // of shared data after the barrier is an agent-scope load issued after thread 0 saw
    // the generation move. The workgroup fences keep the compiler from moving memory
    // operations across the barrier. Per-wave heaps stay private to their wave (one CU,
    // one L2) and are never exchanged.
    template <class Hook>
    __device__ __forceinline__ bool sync(Hook hook) const {
        if (bpg == 1) {
            sc.release();
            __syncthreads();
            if (threadIdx.x == 0) hook();
            sc.release();
            __syncthreads();
            sc.acquire();
            return true;
        }
        constexpr int SC = __HIP_MEMORY_SCOPE_AGENT;
        __shared__ uint32_t s_ok;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores and atomics are complete
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t ok = 1;
            uint32_t g = __hip_atomic_load(&tv->bar_gen, __ATOMIC_RELAXED, SC);
            uint32_t a = __hip_atomic_fetch_add(&tv->bar_count, 1u, __ATOMIC_RELAXED, SC);
            if (a == (uint32_t)bpg - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                hook();
                __hip_atomic_store(&tv->bar_count, 0u, __ATOMIC_RELAXED, SC);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_s_waitcnt(0);  // the hook's writes are complete
                __hip_atomic_fetch_add(&tv->bar_gen, 1u, __ATOMIC_RELAXED, SC);
            } else {
                uint32_t spins = 0;
                while (__hip_atomic_load(&tv->bar_gen, __ATOMIC_RELAXED, SC) == g) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > SPIN_LIMIT || __hip_atomic_load(&tv->abort, __ATOMIC_RELAXED, SC)) {
                        __hip_atomic_store(&tv->abort, 1u, __ATOMIC_RELAXED, SC);
                        ok = 0;
                        break;
                    }
                }
            }
            if (__hip_atomic_load(&tv->abort, __ATOMIC_RELAXED, SC)) ok = 0;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            s_ok = ok;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        return s_ok != 0;
    }
};

// ------------------------------------------------------------------ grid problem
// Item key: (fq << cbits) | (x << ybits | y), fq = f_key >> qshift.
// Best word per cell: (g_key << 5) | (open << 4) | dir, dir = move into the cell (8 = start).
struct GridProb {
    static constexpr int MAXS = 8;
    static constexpr bool kContinue = true;
    QueryDesc q;
    uint64_t cmask, ymask;
    uint64_t goal_id;
    __device__ explicit GridProb(const QueryDesc &qd) : q(qd) {
        cmask = (1ull << q.cbits) - 1;
        ymask = (1ull << q.ybits) - 1;
        goal_id = (uint64_t)q.ex * (uint64_t)q.W + (uint64_t)q.ey;
    }
    __device__ uint64_t item(uint64_t fkey, int x, int y) const {
        return ((fkey >> q.qshift) << q.cbits) | ((uint64_t)x << q.ybits) | (uint64_t)y;
    }
    __device__ bool trivial() const { return q.sx == q.ex && q.sy == q.ey; }
    // Called by one lane: writes the start record (g 0, direction 8, the cell's open bit
    // kept) and returns its item key.
    __device__ uint64_t seed(const Scope &sc) const {
        uint64_t id = (uint64_t)q.sx * q.W + q.sy;
        sc.store(&q.best[id], (uint64_t)((sc.load(&q.best[id]) & kOpenBit) | 8u));
        return item(octile_key(q.sx, q.sy, q.ex, q.ey), q.sx, q.sy);
    }
    __device__ bool goal_done(uint64_t goal, uint64_t lb) const {
        return goal != SENT && goal <= ((lb >> q.cbits) << q.qshift);
    }
    // A popped key whose f is at least the best goal found so far cannot lead to a cheaper
    // goal (with a consistent h every descendant's f is at least its own): it is dropped
    // unexpanded. The key's f field rounds down, so the test never drops a cheaper key.
    __device__ __forceinline__ bool dominated(uint64_t it, uint64_t best_goal) const {
        return best_goal != SENT && ((it >> q.cbits) << q.qshift) >= best_goal;
    }
    // Octile key from the absolute offsets to the goal (each below 2^31).
    __device__ __forceinline__ static uint64_t oct(uint32_t adx, uint32_t ady) {
        uint32_t mn = adx < ady ? adx : ady, mx = adx < ady ? ady : adx;
        return ((uint64_t)(mx - mn) << 36) + (uint64_t)mn * KEY_SQRT2;
    }
    __device__ __forceinline__ static uint32_t uabs(int v) { return (uint32_t)(v < 0 ? -v : v); }

    // Per-lane expansion of one popped item; successors land in s[0..7] (SENT = none).
    // A best word is (g_key << 5) | (open << 4) | dir: it carries its cell's open bit, so
    // the expansion needs no bitmap. Memory is touched in three waves of independent
    // accesses: the cell's best word; every in-range neighbour's best word; an atomicMin
    // for each open neighbour (or the goal) the move would improve. Coordinates and
    // offsets are 32-bit; a neighbour's word is addressed from the cell's (rows +-W).
    // *gk_out receives the cell's g key when the cell was expanded. A key whose exact f
    // (the cell's g key plus its octile key) is at least goal_bound, the best goal found
    // so far, is dropped unexpanded: with a consistent h no descendant reaches a cheaper
    // goal. (The key's own f field is rounded down, so only this exact test can drop the
    // keys of the goal's own f.)
    template <class SC>
    __device__ __forceinline__ void expand(const SC &sc, uint64_t it, uint64_t (&s)[MAXS], int &cnt,
                                           uint64_t &goalmin, uint32_t &err, Ctr &k,
                                           uint64_t *gk_out = nullptr, uint64_t goal_bound = SENT) const {
        const uint64_t cell = it & cmask;
        const int x = (int)(cell >> q.ybits), y = (int)(cell & ymask);
        if (x >= q.H || y >= q.W) {  // never produced by a correct round: fail the query, touch nothing
            err |= FLAG_ABORT;
            return;
        }
        UA_GLOBAL uint64_t *const bp = gptr(q.best) + ((uint64_t)(uint32_t)x * (uint32_t)q.W + (uint32_t)y);
        const uint64_t bw = sc.load(bp);  // the stale test needs the word that pushed the key, or a lower one
        const uint64_t gk = bw >> kGridKeyShift;
        const uint64_t fk = gk + oct(uabs(x - q.ex), uabs(y - q.ey));
        if ((fk >> q.qshift) != (it >> q.cbits)) return;  // stale
        if (fk >= goal_bound) return;                     // cannot lead to a cheaper goal
        ++k.exp;
        if (!(bw & kOpenBit)) return;  // source-cell mask rule: a blocked cell has no moves
        if (gk_out) *gk_out = gk;
        relax(sc, x, y, gk, bp, s, cnt, goalmin, err, k);
    }

    // Continuation inside the round: when the only successors a lane inserted are a child
    // whose f equals its parent's (same f field of the key) and at most one other, the
    // lane expands that child at once instead of pushing it. Its g is the one this lane's
    // atomicMin just wrote (gk + the move's cost); if another lane lowered it since, that
    // lane pushed the child with the lower g and this expansion only produced dominated
    // relaxations. The child is never the goal (the goal is not pushed) and is open (only
    // open cells are pushed). Its successors exclude the parent, so the lane adds at most
    // 1 + 7 keys to the round's buffer, within the 8 per pop it is sized for.
    // Returns the child's key with *dir its move, or SENT.
    __device__ __forceinline__ uint64_t pick_child(uint64_t it, const uint64_t (&s)[MAXS], int cnt, int *dir) const {
        uint64_t best = SENT;
        int bd = 0;
#pragma unroll
        for (int d = 0; d < MAXS; ++d)
            if (s[d] < best) {
                best = s[d];
                bd = d;
            }
        *dir = bd;
        return (cnt >= 1 && cnt <= 2 && (best >> q.cbits) == (it >> q.cbits)) ? best : SENT;
    }
    template <class SC>
    __device__ __forceinline__ void expand_child(const SC &sc, uint64_t child, uint64_t gk, uint64_t (&s)[MAXS],
                                                 int &cnt, uint64_t &goalmin, uint32_t &err, Ctr &k) const {
        const uint64_t cell = child & cmask;
        const int x = (int)(cell >> q.ybits), y = (int)(cell & ymask);
        UA_GLOBAL uint64_t *const bp = gptr(q.best) + ((uint64_t)(uint32_t)x * (uint32_t)q.W + (uint32_t)y);
        ++k.exp;
        relax(sc, x, y, gk, bp, s, cnt, goalmin, err, k);
    }

    // Relaxes the moves out of open cell (x, y) of g key gk (bp its best word).
    template <class SC>
    __device__ __forceinline__ void relax(const SC &sc, int x, int y, uint64_t gk, UA_GLOBAL uint64_t *bp,
                                          uint64_t (&s)[MAXS],
                                          int &cnt, uint64_t &goalmin, uint32_t &err, Ctr &k) const {
        const int gx = x - q.ex, gy = y - q.ey;  // offsets to the goal
        const int64_t Wl = q.W;
        uint64_t w[8];
        uint32_t live = 0;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            const int nx = x + kDX[d], ny = y + kDY[d];
            if ((uint32_t)nx >= (uint32_t)q.H || (uint32_t)ny >= (uint32_t)q.W) continue;
            live |= 1u << d;
            w[d] = sc.load_weak(bp + (kDX[d] * Wl + kDY[d]));
        }
        // legal moves: open targets, and the goal even if blocked (any other blocked cell
        // is a dead end); relax with an atomic only where the new cost is lower than what
        // was read (best words only ever decrease, so a skipped atomic could not succeed).
        // The atomics go out four at a time (orthogonal moves, then diagonal ones), each
        // group issued before any of its results is used, so that they are in flight
        // together; old[j] = 0 marks a move without an atomic.
#pragma unroll
        for (int d0 = 0; d0 < 8; d0 += 4) {
            uint64_t old[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int d = d0 + j;
                old[j] = 0;
                if (!((live >> d) & 1u)) continue;
                const bool is_goal = gx + kDX[d] == 0 && gy + kDY[d] == 0;
                if (!is_goal && !(w[d] & kOpenBit)) continue;
                ++k.gen;
                const uint64_t ng = gk + (d < 4 ? KEY_ONE : KEY_SQRT2);
                if ((w[d] >> kGridKeyShift) <= ng) continue;  // no key: its f is not needed
                const uint64_t nf = ng + oct(uabs(gx + kDX[d]), uabs(gy + kDY[d]));
                if (nf >= KEY_LIMIT) {
                    err |= FLAG_RANGE;
                    continue;
                }
                ++k.atom;
                old[j] = sc.fmin(bp + (kDX[d] * Wl + kDY[d]), (ng << kGridKeyShift) | (w[d] & kOpenBit) | (uint64_t)d);
                if (!is_goal)
                    s[d] = ((nf >> q.qshift) << q.cbits) | ((uint64_t)(uint32_t)(x + kDX[d]) << q.ybits) |
                           (uint64_t)(uint32_t)(y + kDY[d]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int d = d0 + j;
                const uint64_t ng = gk + (d < 4 ? KEY_ONE : KEY_SQRT2);
                if ((old[j] >> kGridKeyShift) <= ng) {  // no atomic, or it did not improve the cell
                    s[d] = SENT;
                    continue;
                }
                ++k.ins;
                if (gx + kDX[d] == 0 && gy + kDY[d] == 0) goalmin = goalmin < ng ? goalmin : ng;
                else ++cnt;
            }
        }
    }
};

// ------------------------------------------------------------------ puzzle problem
// Board: 4 bits per cell, cell i at bits 4i..4i+3 (PuzzleStorage<4>, puzzle/storage.hpp:86-139;
// N = 3 uses the same layout with 9 cells). Pattern databases as in
// puzzle/GPU-kernel.cuh:71-135: rank of the k-tuple of tile cells, one table per partition.
template <int N>
struct PSpec;
template <>
struct PSpec<3> {
    static constexpr int NP = 1;
    static constexpr int K[1] = {8};
    static constexpr int T[1][8] = {{1, 2, 3, 4, 5, 6, 7, 8}};
    static constexpr uint32_t M[1][8] = {{40320, 5040, 720, 120, 24, 6, 2, 1}};
    static constexpr uint64_t OFF[1] = {0};
};
template <>
struct PSpec<4> {
    static constexpr int NP = 2;
    static constexpr int K[2] = {8, 7};
    static constexpr int T[2][8] = {{1, 2, 3, 4, 5, 6, 9, 13}, {7, 8, 10, 11, 12, 14, 15, 0}};
    static constexpr uint32_t M[2][8] = {{32432400, 2162160, 154440, 11880, 990, 90, 9, 1},
                                         {3603600, 240240, 17160, 1320, 110, 10, 1, 0}};
    static constexpr uint64_t OFF[2] = {0, 518918400ull};
};

template <>
struct PSpec<5> {
    static constexpr int NP = 4;
    static constexpr int K[4] = {6, 6, 6, 6};
    static constexpr int T[4][8] = {{3, 4, 5, 10, 15, 20, 0, 0},
                                    {2, 1, 6, 11, 16, 21, 0, 0},
                                    {7, 8, 9, 12, 17, 22, 0, 0},
                                    {13, 14, 18, 19, 23, 24, 0, 0}};
    static constexpr uint32_t M[4][8] = {{5100480, 212520, 9240, 420, 20, 1, 0, 0},
                                         {5100480, 212520, 9240, 420, 20, 1, 0, 0},
                                         {5100480, 212520, 9240, 420, 20, 1, 0, 0},
                                         {5100480, 212520, 9240, 420, 20, 1, 0, 0}};
    static constexpr uint64_t OFF[4] = {0, 127512000ull, 255024000ull, 382536000ull};
};

__device__ __forceinline__ int nib_pos(uint64_t s, int t) {
    uint64_t x = s ^ (0x1111111111111111ull * (uint64_t)t);
    uint64_t z = (x - 0x1111111111111111ull) & ~x & 0x8888888888888888ull;
    return __builtin_ctzll(z) >> 2;
}
__device__ __forceinline__ uint64_t goal_board(int n) {
    uint64_t g = 0;
    for (int i = 0; i < n * n - 1; ++i) g |= (uint64_t)(i + 1) << (4 * i);
    return g;
}

template <int N>
__device__ __forceinline__ uint32_t pattern_rank(uint64_t s, int p) {
    using S = PSpec<N>;
    uint32_t mask = 0, code = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < S::K[p]) {
            int pos = nib_pos(s, S::T[p][j]);
            uint32_t v = (uint32_t)pos - __builtin_popcount(mask & ((1u << pos) - 1u));
            code += v * S::M[p][j];
            mask |= 1u << pos;
        }
    }
    return code;
}

// The partitions of PSpec<3> / PSpec<4> as one table indexed at run time, so that a
// single copy of the search serves both board sizes.
struct PTab {
    int np;             // partitions
    int k[2];           // tiles per partition
    int t[2][8];        // the tiles
    uint32_t m[2][8];   // rank radices
    uint64_t off[2];    // byte offset of each partition's table in the PDB
};
__constant__ static const PTab kPTab[2] = {
    {PSpec<3>::NP, {PSpec<3>::K[0], 0},
     {{PSpec<3>::T[0][0], PSpec<3>::T[0][1], PSpec<3>::T[0][2], PSpec<3>::T[0][3], PSpec<3>::T[0][4],
       PSpec<3>::T[0][5], PSpec<3>::T[0][6], PSpec<3>::T[0][7]},
      {0, 0, 0, 0, 0, 0, 0, 0}},
     {{PSpec<3>::M[0][0], PSpec<3>::M[0][1], PSpec<3>::M[0][2], PSpec<3>::M[0][3], PSpec<3>::M[0][4],
       PSpec<3>::M[0][5], PSpec<3>::M[0][6], PSpec<3>::M[0][7]},
      {0, 0, 0, 0, 0, 0, 0, 0}},
     {PSpec<3>::OFF[0], 0}},
    {PSpec<4>::NP, {PSpec<4>::K[0], PSpec<4>::K[1]},
     {{PSpec<4>::T[0][0], PSpec<4>::T[0][1], PSpec<4>::T[0][2], PSpec<4>::T[0][3], PSpec<4>::T[0][4],
       PSpec<4>::T[0][5], PSpec<4>::T[0][6], PSpec<4>::T[0][7]},
      {PSpec<4>::T[1][0], PSpec<4>::T[1][1], PSpec<4>::T[1][2], PSpec<4>::T[1][3], PSpec<4>::T[1][4],
       PSpec<4>::T[1][5], PSpec<4>::T[1][6], PSpec<4>::T[1][7]}},
     {{PSpec<4>::M[0][0], PSpec<4>::M[0][1], PSpec<4>::M[0][2], PSpec<4>::M[0][3], PSpec<4>::M[0][4],
       PSpec<4>::M[0][5], PSpec<4>::M[0][6], PSpec<4>::M[0][7]},
      {PSpec<4>::M[1][0], PSpec<4>::M[1][1], PSpec<4>::M[1][2], PSpec<4>::M[1][3], PSpec<4>::M[1][4],
       PSpec<4>::M[1][5], PSpec<4>::M[1][6], PSpec<4>::M[1][7]}},
     {PSpec<4>::OFF[0], PSpec<4>::OFF[1]}},
};

__device__ __forceinline__ uint32_t pattern_rank_rt(const PTab &tb, uint64_t s, int p) {
    uint32_t mask = 0, code = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < tb.k[p]) {
            int pos = nib_pos(s, tb.t[p][j]);
            uint32_t v = (uint32_t)pos - __builtin_popcount(mask & ((1u << pos) - 1u));
            code += v * tb.m[p][j];
            mask |= 1u << pos;
        }
    }
    return code;
}

__device__ __forceinline__ int puzzle_h(const PTab &tb, uint64_t s, const uint8_t *pdb) {
    int h = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
        if (p < tb.np) h += gptr(pdb)[tb.off[p] + pattern_rank_rt(tb, s, p)];
    return h;
}

// Item key: (f << 56) | ((255 - g) << 48) | (h1 << 40) | slot (ties: larger g first), h1
// the state's second-partition PDB term (0 for N = 3). With the state's h = f - g (f - g >>
// wshift when weighted) known from the key, a child's h needs one gather: a move shifts one
// tile, so only that tile's partition term changes.
// Hash: slot s is 16 bytes, best[2 s] = board (0 = empty) and the low half of best[2 s + 1]
// = ~((g << 4) | dir) (dir 15 = start). The value is kept complemented so that a zeroed
// table is empty (0 = no value) and lowering it is an atomicMax; the expansion reads a
// popped state's board and value from one line.
// The side n (3 or 4) is a run-time value.
constexpr uint32_t tile_mask(const int (&t)[8], int k) {
    uint32_t m = 0;
    for (int j = 0; j < k; ++j) m |= 1u << t[j];
    return m;
}
constexpr uint32_t kPart1Mask4 = tile_mask(PSpec<4>::T[1], PSpec<4>::K[1]);  // tiles of the 15-puzzle's second partition
struct PuzzleProb {
    static constexpr int MAXS = 4;
    static constexpr bool kContinue = false;
    static constexpr int MAXPROBE = 256;
    QueryDesc q;
    const uint8_t *pdb;
    const PTab &tb;
    int n;
    uint64_t hmask;
    uint64_t goal_b;
    uint32_t p1mask;  // tiles of the second partition (none for N = 3)
    __device__ __forceinline__ PuzzleProb(const QueryDesc &qd, const uint8_t *p) : q(qd), pdb(p), tb(kPTab[qd.n == 3 ? 0 : 1]) {
        n = qd.n == 3 ? 3 : 4;
        hmask = (1ull << q.hash_log2) - 1;
        goal_b = goal_board(n);
        p1mask = n == 4 ? kPart1Mask4 : 0u;
    }
    __device__ bool trivial() const { return q.start_state == goal_b; }
    __device__ __forceinline__ UA_GLOBAL uint64_t *board_word(uint64_t slot) const { return gptr(q.best) + 2 * slot; }
    __device__ __forceinline__ UA_GLOBAL uint32_t *value_word(uint64_t slot) const {
        return reinterpret_cast<UA_GLOBAL uint32_t *>(gptr(q.best) + (2 * slot + 1));
    }
    template <class SC>
    __device__ __forceinline__ uint64_t find_or_insert(const SC &sc, uint64_t b, uint32_t &err,
                                                       uint32_t &atom) const {
        uint64_t slot = mix64(b) & hmask;
        for (int i = 0; i < MAXPROBE; ++i) {
            uint64_t k = sc.load_weak(board_word(slot));
            if (k == b) return slot;
            if (k == 0) {
                uint64_t exp = 0;
                ++atom;
                if (sc.cas(board_word(slot), exp, b) || exp == b) return slot;
            }
            slot = (slot + 1) & hmask;
        }
        err |= q.bounded ? FLAG_PRUNED : FLAG_CAPACITY;
        return SENT;
    }
    __device__ uint64_t seed(const Scope &sc) const {
        uint32_t err = 0, atom = 0;
        uint64_t slot = find_or_insert(sc, q.start_state, err, atom);
        sc.store(value_word(slot), ~(uint32_t)15);
        const uint64_t h = (uint64_t)puzzle_h(tb, q.start_state, pdb);
        const uint64_t h1 = tb.np > 1 ? gptr(pdb)[tb.off[1] + pattern_rank_rt(tb, q.start_state, 1)] : 0u;
        return (h << 56) | (255ull << 48) | (h1 << 40) | slot;
    }
    // a weighted search (memory-bounded scheme) stops at its first goal
    __device__ bool goal_done(uint64_t goal, uint64_t lb) const {
        return goal != SENT && (q.wshift > 0 || (goal >> 32) <= (lb >> 56));
    }
    // a popped key of f at least the best goal's moves leads to no shorter solution (the
    // PDB heuristic is consistent): dropped unexpanded
    __device__ __forceinline__ bool dominated(uint64_t it, uint64_t best_goal) const {
        return best_goal != SENT && q.wshift == 0 && (it >> 56) >= (best_goal >> 32);
    }
    template <class SC>
    __device__ __forceinline__ void expand(const SC &sc, uint64_t it, uint64_t (&s)[MAXS], int &cnt,
                                           uint64_t &goalmin, uint32_t &err, Ctr &k) const {
        uint64_t slot = it & ((1ull << kPuzzleSlotBits) - 1);
        if (slot > hmask) {  // never produced by a correct round: fail the query, touch nothing
            err |= FLAG_ABORT;
            return;
        }
        uint32_t gi = 255u - (uint32_t)((it >> 48) & 255u);
        const int hp = (int)(it >> 56) - (int)(gi >> q.wshift);  // the state's h and its second term
        const int h1p = (int)((it >> 40) & 255u);
        // the popped state's board and value may have been written from another block:
        // scoped loads (the stale test must see the value that pushed the key, or a lower one)
        uint64_t b = sc.load(board_word(slot));
        uint32_t v = ~sc.load(value_word(slot));
        uint32_t gb = v >> 4;
        if (gb != gi) return;  // stale
        ++k.exp;
        const int dprev = (int)(v & 15u);
        const int bl = nib_pos(b, 0);
        const int bx = n == 4 ? bl >> 2 : bl / 3, by = bl - bx * n;
        const uint32_t ng = gb + 1;
        // The moves are handled loop by loop, so that the PDB gathers of all children are in
        // flight together, and so are the atomicMins on their value words; the hash probes
        // (a load, then a CAS on an empty slot) stay one child at a time.
        uint64_t ns[4];
        int h[4], h1[4];
        uint32_t legal = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            ns[d] = 0;
            h[d] = 0;
            if (dprev < 4 && d == (dprev ^ 1)) continue;  // the move back to the parent
            const int nx = bx + kDX[d], ny = by + kDY[d];
            if (nx < 0 || nx >= n || ny < 0 || ny >= n) continue;
            legal |= 1u << d;
            const int nb = nx * n + ny;
            const uint64_t t = (b >> (4 * nb)) & 15ull;
            ns[d] = b + (t << (4 * bl)) - (t << (4 * nb));
            // the moved tile's partition term is the only one to change: one gather (both
            // ranks are computed, so that the radix tables stay wave-uniform)
            const bool second = (p1mask >> t) & 1u;
            const uint32_t r0 = pattern_rank_rt(tb, ns[d], 0), r1 = pattern_rank_rt(tb, ns[d], 1);
            const int term = gptr(pdb)[second ? tb.off[1] + r1 : tb.off[0] + r0];
            h[d] = second ? hp - h1p + term : term + h1p;
            h1[d] = second ? term : h1p;
        }
        uint64_t ts[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            ts[d] = SENT;
            if (!((legal >> d) & 1u)) continue;
            ++k.gen;
            const uint32_t f = (ng >> q.wshift) + (uint32_t)h[d];
            if (f > 255u) {
                err |= q.bounded ? FLAG_PRUNED : FLAG_RANGE;
                continue;
            }
            ts[d] = find_or_insert(sc, ns[d], err, k.atom);
        }
        uint32_t old[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            old[d] = 0;  // no atomic: never an improvement
            if (ts[d] != SENT) {
                ++k.atom;
                old[d] = ~sc.fmax(value_word(ts[d]), ~((ng << 4) | (uint32_t)d));
            }
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            if (ts[d] == SENT || (old[d] >> 4) <= ng) continue;
            ++k.ins;
            if (ns[d] == goal_b) {
                const uint64_t gk = ((uint64_t)ng << 32) | ts[d];
                goalmin = goalmin < gk ? goalmin : gk;
            } else {
                const uint32_t f = (ng >> q.wshift) + (uint32_t)h[d];
                s[d] = ((uint64_t)f << 56) | ((uint64_t)(255u - ng) << 48) | ((uint64_t)h1[d] << 40) | ts[d];
                ++cnt;
            }
        }
    }
};

// ------------------------------------------------------------------ 5x5 puzzle problem
// Board of two words as PuzzleStorage<5> (puzzle/storage.hpp:141-217): cells 0..11 at
// w0 bits 5c..5c+4, cells 12..23 at w1 bits 5(c-12).., cell 24 split into its low four
// bits at w0 bits 60..63 and its bit 4 at w1 bit 60. Neither word of a board is zero
// (each holds at least eleven nonzero tiles), so 0 marks an empty hash word.
constexpr uint64_t kOnes12 = 0x84210842108421ull;   // bit 5i, i < 12
constexpr uint64_t kLow60 = (1ull << 60) - 1;

__device__ __forceinline__ uint64_t cell5(uint64_t w0, uint64_t w1, int c) {
    if (c < 12) return (w0 >> (5 * c)) & 31;
    if (c < 24) return (w1 >> (5 * (c - 12))) & 31;
    return (w0 >> 60) | ((w1 >> 56) & 16);
}
// toggles tile t in cell c (sets it in an empty cell, clears it where it is)
__device__ __forceinline__ void xor5(uint64_t &w0, uint64_t &w1, int c, uint64_t t) {
    if (c < 12) w0 ^= t << (5 * c);
    else if (c < 24) w1 ^= t << (5 * (c - 12));
    else {
        w0 ^= (t & 15) << 60;
        w1 ^= (t & 16) << 56;
    }
}
// cell of tile t: the lowest zero field of (cells ^ t) in either word, else cell 24
__device__ __forceinline__ int pos5(uint64_t w0, uint64_t w1, int t) {
    const uint64_t pat = kOnes12 * (uint64_t)t;
    const uint64_t x0 = (w0 & kLow60) ^ pat, x1 = (w1 & kLow60) ^ pat;
    const uint64_t z0 = (x0 - kOnes12) & ~x0 & (kOnes12 << 4);
    const uint64_t z1 = (x1 - kOnes12) & ~x1 & (kOnes12 << 4);
    if (z0) return __builtin_ctzll(z0) / 5;
    if (z1) return 12 + __builtin_ctzll(z1) / 5;
    return 24;
}
__device__ __forceinline__ uint32_t rank5(uint64_t w0, uint64_t w1, int p) {
    using S = PSpec<5>;
    uint32_t mask = 0, code = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        int pos = pos5(w0, w1, S::T[p][j]);
        code += ((uint32_t)pos - __builtin_popcount(mask & ((1u << pos) - 1u))) * S::M[p][j];
        mask |= 1u << pos;
    }
    return code;
}
// partition of each tile (tile 0, the blank, in none)
__constant__ static const int8_t kPart5[25] = {-1, 1, 1, 0, 0, 0, 1, 2, 2, 2, 0, 1, 2,
                                               3, 3, 0, 1, 2, 3, 3, 0, 1, 2, 3, 3};

// Item key: (f << 56) | ((255 - g) << 48) | (h << 40) | slot; hash values as
// PuzzleProb; hash keys are two words per slot (keys[2 slot], keys[2 slot + 1]). A slot is claimed by a CAS of its first word, and
// the winner stores the second word at once; a prober that finds its first word in a
// slot whose second word is not yet stored reads it again (a bounded wait). The child's
// h is the parent's (from its item) with the moved tile's partition term swapped.
struct PuzzleProb5 {
    static constexpr int MAXS = 4;
    static constexpr bool kContinue = false;
    static constexpr int MAXPROBE = 256;
    static constexpr int SPIN = 1 << 16;
    QueryDesc q;
    uint64_t hmask;
    uint64_t g0, g1;
    __device__ explicit PuzzleProb5(const QueryDesc &qd) : q(qd) {
        hmask = (1ull << q.hash_log2) - 1;
        g0 = 0x62d4941cc520c41ull;  // tiles 1..12 on cells 0..11
        g1 = 0xc5ed5a4e5183dcdull;  // tiles 13..24 on cells 12..23, blank on cell 24
    }
    __device__ bool trivial() const { return q.start_state == g0 && q.start_state2 == g1; }
    __device__ __forceinline__ int h_full(uint64_t w0, uint64_t w1) const {
        int h = 0;
#pragma unroll
        for (int p = 0; p < 4; ++p) h += gptr(q.pdb)[PSpec<5>::OFF[p] + rank5(w0, w1, p)];
        return h;
    }
    template <class SC>
    __device__ __forceinline__ uint64_t find_or_insert(const SC &sc, uint64_t b0, uint64_t b1, uint32_t &err,
                                                       uint32_t &atom) const {
        uint64_t slot = mix64(b0 ^ mix64(b1)) & hmask;
        for (int i = 0; i < MAXPROBE; ++i) {
            uint64_t k0 = sc.load_weak(gptr(q.best) + (2 * slot));
            if (k0 == 0) {
                uint64_t exp = 0;
                ++atom;
                if (sc.cas(gptr(q.best) + (2 * slot), exp, b0)) {
                    sc.store(gptr(q.best) + (2 * slot + 1), b1);
                    return slot;
                }
                k0 = exp;
            }
            if (k0 == b0) {
                uint64_t k1 = sc.load(gptr(q.best) + (2 * slot + 1));
                for (int w = 0; k1 == 0 && w < SPIN; ++w) k1 = sc.load(gptr(q.best) + (2 * slot + 1));
                if (k1 == 0) {
                    err |= FLAG_ABORT;
                    return SENT;
                }
                if (k1 == b1) return slot;
            }
            slot = (slot + 1) & hmask;
        }
        err |= q.bounded ? FLAG_PRUNED : FLAG_CAPACITY;
        return SENT;
    }
    __device__ uint64_t seed(const Scope &sc) const {
        uint32_t err = 0, atom = 0;
        uint64_t slot = find_or_insert(sc, q.start_state, q.start_state2, err, atom);
        if (slot == SENT) return SENT;
        sc.store(gptr(q.vals) + (slot), (uint32_t)15);
        uint64_t h = (uint64_t)h_full(q.start_state, q.start_state2);
        return (h << 56) | (255ull << 48) | (h << 40) | slot;
    }
    // a weighted search (memory-bounded scheme) stops at its first goal
    __device__ bool goal_done(uint64_t goal, uint64_t lb) const {
        return goal != SENT && (q.wshift > 0 || (goal >> 32) <= (lb >> 56));
    }
    // a popped key of f at least the best goal's moves leads to no shorter solution (the
    // PDB heuristic is consistent): dropped unexpanded
    __device__ __forceinline__ bool dominated(uint64_t it, uint64_t best_goal) const {
        return best_goal != SENT && q.wshift == 0 && (it >> 56) >= (best_goal >> 32);
    }
    template <class SC>
    __device__ __forceinline__ void expand(const SC &sc, uint64_t it, uint64_t (&s)[MAXS], int &cnt,
                                           uint64_t &goalmin, uint32_t &err, Ctr &k) const {
        uint64_t slot = it & ((1ull << kPuzzleSlotBits) - 1);
        if (slot > hmask) {  // never produced by a correct round: fail the query, touch nothing
            err |= FLAG_ABORT;
            return;
        }
        uint32_t gi = 255u - (uint32_t)((it >> 48) & 255u);
        int hp = (int)((it >> 40) & 255u);
        // written from any block of the team: scoped loads (see PuzzleProb::expand)
        uint64_t w0 = sc.load(gptr(q.best) + (2 * slot));
        uint64_t w1 = sc.load(gptr(q.best) + (2 * slot + 1));
        uint32_t v = sc.load(gptr(q.vals) + (slot));
        uint32_t gb = v >> 4;
        if (gb != gi) return;  // stale
        ++k.exp;
        int dprev = (int)(v & 15u);
        int bl = pos5(w0, w1, 0);
        int bx = bl / 5, by = bl - bx * 5;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            if (dprev < 4 && d == (dprev ^ 1)) continue;  // the move back to the parent
            int nx = bx + kDX[d], ny = by + kDY[d];
            if (nx < 0 || nx >= 5 || ny < 0 || ny >= 5) continue;
            ++k.gen;
            int nb = nx * 5 + ny;
            uint64_t t = cell5(w0, w1, nb);
            uint64_t n0 = w0, n1 = w1;
            xor5(n0, n1, bl, t);
            xor5(n0, n1, nb, t);
            const int p = kPart5[t];
            const uint64_t off = (uint64_t)p * 127512000ull;  // PSpec<5>::OFF[p]
            int h = hp - (int)gptr(q.pdb)[off + rank5(w0, w1, p)] + (int)gptr(q.pdb)[off + rank5(n0, n1, p)];
            uint32_t ng = gb + 1;
            const uint32_t f = (ng >> q.wshift) + (uint32_t)h;
            if (f > 255u) {
                err |= q.bounded ? FLAG_PRUNED : FLAG_RANGE;
                continue;
            }
            uint64_t ns_slot = find_or_insert(sc, n0, n1, err, k.atom);
            if (ns_slot == SENT) continue;
            ++k.atom;
            uint32_t old = sc.fmin(gptr(q.vals) + (ns_slot), (ng << 4) | (uint32_t)d);
            if ((old >> 4) <= ng) continue;
            ++k.ins;
            if (n0 == g0 && n1 == g1) {
                uint64_t gk = ((uint64_t)ng << 32) | ns_slot;
                goalmin = goalmin < gk ? goalmin : gk;
            } else {
                s[d] = ((uint64_t)f << 56) | ((uint64_t)(255u - ng) << 48) | ((uint64_t)h << 40) | ns_slot;
                ++cnt;
            }
        }
    }
};

// ------------------------------------------------------------------ the team search
template <class Team, class Prob>
__device__ __forceinline__ void team_solve(const QueryDesc &q, QueryResult *res, const Team &team, const Prob &P) {
    const Scope sc = team.sc;
    const int l = lane_id();
    const int tw = team.tw, W = team.W;
    RoundSlot *const slots = team.m.slot;
    uint64_t *const stats = team.m.stats;
    const int segcap = team.m.segcap;

    if (P.trivial()) {  // start == goal: cost 0, no search
        if (team.leader()) {
            res->status = ST_OK;
            res->found = 1;
            res->goal = 0;
            res->rounds = 0;
            res->pops = res->expanded = res->generated = res->inserted = res->atomics = 0;
            res->flags = 0;
        }
        return;
    }

    WaveHeap heap;
    heap.init(q.heap + (size_t)tw * (size_t)(q.heap_cap - 1) * 64, q.heap_cap);
    // a query suspended by an earlier serving launch resumes from its record
    const bool resumed = q.resume != nullptr && q.resume[0] != 0;
    uint64_t *rec_heap = q.resume ? q.resume + RESUME_HEADER + (size_t)tw * RESUME_PER_HEAP : nullptr;
    if (team.leader()) {
        for (int i = 0; i < 2; ++i) {
            sc.store(&slots[i].lb, SENT);
            sc.store(&slots[i].goal, SENT);
            sc.store(&slots[i].cnt, 0u);
            sc.store(&slots[i].err, 0u);
        }
        for (int i = 0; i < 4; ++i) sc.store(&stats[i], resumed ? q.resume[3 + i] : (uint64_t)0);
        sc.store(&stats[4], resumed ? q.resume[8] : (uint64_t)0);
    }
    if (resumed) {
        heap.root = rec_heap[l];
        heap.part = rec_heap[64 + l];
        heap.n = (int)rec_heap[128];
        heap.pc = (int)rec_heap[129];
    } else if (tw == 0) {
        uint64_t it = SENT;
        if (l == 0) it = P.seed(sc);
        heap.part = it;
        heap.pc = 1;
    }
    if (!team.sync([] {})) return;

    uint32_t n_pop = 0;
    Ctr k;
    uint64_t best_goal = resumed ? q.resume[2] : SENT;
    uint32_t pruned = resumed ? (uint32_t)q.resume[7] & FLAG_PRUNED : 0u;
    int32_t status = ST_OK;
    int found = 0;
    const int r0 = resumed ? (int)q.resume[1] : 0;
    int r = r0;
    for (;; ++r) {
        RoundSlot *cur = &slots[r & 1];
        uint64_t *cur_seg = team.m.seg + (size_t)(r & 1) * segcap;
        // ---- phase A: pop, expand, relax, append
        uint64_t it = heap.pop();
        uint64_t wmin = readlane64(it, 0);
        if (l == 0 && wmin != SENT) sc.fmin(&cur->lb, wmin);

        uint64_t s[Prob::MAXS];
#pragma unroll
        for (int i = 0; i < Prob::MAXS; ++i) s[i] = SENT;
        int c = 0;
        uint64_t goalmin = SENT;
        uint32_t err = 0;
        uint64_t gk = 0;
        if (it != SENT) {
            ++n_pop;
            if (P.dominated(it, best_goal)) it = SENT;
        }
        if (it != SENT) {
            if constexpr (Prob::kContinue) P.expand(sc, it, s, c, goalmin, err, k, &gk);
            else P.expand(sc, it, s, c, goalmin, err, k);
        }
        // appends the lane's keys -- s[] and `keep` unless SENT, c in all -- to the round's
        // insert buffer
        uint64_t keep = SENT;
        auto append = [&] {
            int total;
            const int off = wave_excl_scan(c, &total);
            uint32_t base = 0;
            if (l == 0 && total) base = sc.fadd(&cur->cnt, (uint32_t)total);
            base = __shfl(base, 0);
            uint64_t *out = cur_seg + base + off;  // shared by the team's blocks: scoped stores
#pragma unroll
            for (int i = 0; i < Prob::MAXS; ++i)
                if (s[i] != SENT) sc.store(out++, s[i]);
            if (keep != SENT) sc.store(out, keep);
        };
        if constexpr (Prob::kContinue) {
            // Only a wave whose heap still holds keys after its pop continues: its pops were
            // the smallest of more keys than a round takes, so best-first order still holds.
            // When the open list fits in one round's pops, continuation would expand children
            // on g values that later rounds correct, and their re-expansions multiply.
            int dc = 0;
            const uint64_t child = it != SENT && heap.size() > 0 ? P.pick_child(it, s, c, &dc) : SENT;
            if (__ballot(child != SENT)) {
                if (child != SENT) {  // the child's sibling, if any, waits in `keep`
#pragma unroll
                    for (int i = 0; i < Prob::MAXS; ++i) {
                        if (i != dc && s[i] < keep) keep = s[i];
                        s[i] = SENT;
                    }
                    c = keep != SENT;
                    P.expand_child(sc, child, gk + (dc < 4 ? KEY_ONE : KEY_SQRT2), s, c, goalmin, err, k);
                }
            }
        }
        append();
        if (__ballot(goalmin != SENT)) {  // rare: reduce only in rounds that reach a goal
            goalmin = wave_min64(goalmin);
            if (l == 0) sc.fmin(&cur->goal, goalmin);
        }
        if (__ballot(err != 0)) {
            for (int j = 32; j > 0; j >>= 1) err |= __shfl_xor(err, j);
            if (l == 0) sc.fold_or(&cur->err, err);
        }

        RoundSlot *nxt = &slots[(r + 1) & 1];
        TeamVars *const tvp = team.tv;
        bool ok = team.sync([nxt, cur, sc, tvp] {
            sc.store(&nxt->lb, SENT);
            sc.store(&nxt->goal, SENT);
            sc.store(&nxt->cnt, 0u);
            sc.store(&nxt->err, 0u);
            uint64_t dl = __hip_atomic_load(&tvp->deadline, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (dl && (uint64_t)now_ticks() > dl) sc.fold_or(&cur->err, (uint32_t)FLAG_TIMEUP);
        });
        if (!ok) {
            status = ST_ABORT;
            break;
        }
        // ---- phase B: termination test, then insert this wave's share. One lane reads
        // the round slot and hands it to the block through LDS, so every wave of the
        // block takes its decisions from the same values.
        RoundSlot *const view = team.m.view;
        if (threadIdx.x == 0) {
            view->lb = sc.load(&cur->lb);
            view->goal = sc.load(&cur->goal);
            view->cnt = sc.load(&cur->cnt);
            view->err = sc.load(&cur->err);
        }
        __syncthreads();
        const uint64_t lb = view->lb, g = view->goal;
        const uint32_t cnt = view->cnt, rerr = view->err;  // rewritten only after the next team barrier
        best_goal = best_goal < g ? best_goal : g;
        pruned |= rerr & FLAG_PRUNED;
        if (rerr & ~(uint32_t)(FLAG_TIMEUP | FLAG_PRUNED)) {
            status = (rerr & FLAG_ABORT) ? ST_ABORT : (rerr & FLAG_CAPACITY) ? ST_CAPACITY : ST_RANGE;
            break;
        }
        if (lb == SENT) {  // every heap was empty: the open list is exhausted
            found = best_goal != SENT;
            status = found ? ST_OK : ST_NO_PATH;
            break;
        }
        if (P.goal_done(best_goal, lb)) {
            found = 1;
            break;
        }
        if (r + 1 >= q.max_rounds) {
            status = ST_ROUNDS;
            break;
        }
        int rot = (tw + r) % W;
        int mine = cnt > (uint32_t)rot ? (int)((cnt - (uint32_t)rot + W - 1) / W) : 0;
        bool fits = true;
        for (int b0 = 0; b0 < mine; b0 += 64) {
            int k = b0 + l;
            uint64_t v = k < mine ? sc.load(cur_seg + ((size_t)k * W + rot)) : SENT;
            int sreal = mine - b0 < 64 ? mine - b0 : 64;
            v = wave_sort(v);
            fits &= heap.insert(v, sreal);
        }
        if (!fits && l == 0) sc.fold_or(&nxt->err, q.bounded ? (uint32_t)FLAG_PRUNED : (uint32_t)FLAG_CAPACITY);
        if (q.resume != nullptr && ((rerr & FLAG_TIMEUP) || (q.round_budget > 0 && r + 1 - r0 >= q.round_budget))) {
            status = ST_SUSPENDED;  // every wave reaches this with the same r and view
            break;
        }
    }
    if (status == ST_SUSPENDED) {
        rec_heap[l] = heap.root;
        rec_heap[64 + l] = heap.part;
        if (l == 0) {
            rec_heap[128] = (uint64_t)heap.n;
            rec_heap[129] = (uint64_t)heap.pc;
        }
    }

    // ---- statistics and result
    const int64_t t_pop = wave_sum64(n_pop), t_exp = wave_sum64(k.exp), t_gen = wave_sum64(k.gen),
                  t_ins = wave_sum64(k.ins), t_atom = wave_sum64(k.atom);
    if (l == 0) {
        sc.fadd(&stats[0], (uint64_t)t_pop);
        sc.fadd(&stats[1], (uint64_t)t_exp);
        sc.fadd(&stats[2], (uint64_t)t_gen);
        sc.fadd(&stats[3], (uint64_t)t_ins);
        sc.fadd(&stats[4], (uint64_t)t_atom);
    }
    team.sync([] {});
    if (team.leader()) {
        if (status == ST_SUSPENDED) {
            q.resume[1] = (uint64_t)(r + 1);
            q.resume[2] = best_goal;
            for (int i = 0; i < 4; ++i) q.resume[3 + i] = sc.load(&stats[i]);
            q.resume[7] = pruned;
            q.resume[8] = sc.load(&stats[4]);
            q.resume[0] = 1;
        }
        res->status = status;
        res->found = found;
        res->goal = best_goal;
        res->rounds = r + 1;
        res->pops = (int64_t)sc.load(&stats[0]);
        res->expanded = (int64_t)sc.load(&stats[1]);
        res->generated = (int64_t)sc.load(&stats[2]);
        res->inserted = (int64_t)sc.load(&stats[3]) + 1;
        res->atomics = (int64_t)sc.load(&stats[4]);
        res->flags = pruned;
    }
}

// ------------------------------------------------------------------ bucket open list (grids)
// The open list of a grid team as one ring of kBuckets buckets shared by the team, instead
// of one sorted heap per wave. A key's bucket is its f field (fq = f_key >> qshift) over
// 2^bshift, i.e. f in steps of 2^(qshift + bshift - 36) (1/16 step as planned); its ring
// slot is that bucket mod kBuckets. Each ring bucket holds 2^bkt_log2 keys in HBM as a circular array; per ring
// bucket the team counts the keys appended (tail) and consumed (head). With a consistent
// heuristic a child's f is at least its parent's and at most 2 sqrt(2) steps more, and a
// round takes keys only from the lowest bwin buckets, so every key in the open list lies
// within kBuckets buckets of the lowest one: the ring never aliases (a key outside the
// window fails the query; it cannot occur). No key is compared with another: a round takes
// the lowest buckets' keys in arrival order, which replaces the heaps' sorting networks.
//
// A round:
//   (A) every wave reads the ring's counters in window order (bucket cur + 4 lane + i);
//   barrier (every copy is taken before this round's appends)
//   (B) plan, identical in every wave: the lowest live bucket blo; the team stops when the
//       best goal is at most blo's lower f edge (every open key lies at or above it) or
//       no bucket is live; otherwise it takes up to 64 keys a wave from buckets
//       blo .. blo + bwin - 1, in bucket and arrival order; the team's first wave writes
//       the next round's consumed counts (the other copy of head);
//   (C) lane l of the team's wave w loads the key of rank 64 w + l;
//   (D) expansion and relaxation as on the heaps (GridProb::expand; continuation when the
//       window held more keys than the round takes);
//   (E) each improved successor is appended to its bucket: an atomic add on its tail
//       gives the position, the key goes to HBM;
//   barrier (the appends, the round's goal and errors are complete)
// This is the round of uAstar's kExtractExpand / kDeduplicate / kHeapInsert
// (pathway/GPU-kernel.cuh:169-544) with its termination test on the device every round.
//
// The ring runs on one-block teams only. Its counters live in LDS (copied to the resume
// record when a serving launch suspends the query); the first kBucketMetaWords words of
// the query's bucket storage are reserved for them and stay unused.
struct BucketMeta {
    uint32_t tail[kBuckets];
    uint32_t head[2][kBuckets];  // consumed counts: round r reads head[r & 1], writes head[(r + 1) & 1]
};
static_assert(sizeof(BucketMeta) <= kBucketMetaWords * 8, "bucket counters fit their reserved words");

// Relaxed atomics at workgroup scope fixed at compile time, for the best words of a
// one-block team's grid (the team is their only user).
struct WgScope {
    template <class T>
    __device__ __forceinline__ T load(const UA_GLOBAL T *p) const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <class T>
    __device__ __forceinline__ T load_weak(const UA_GLOBAL T *p) const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <class T>
    __device__ __forceinline__ void store(UA_GLOBAL T *p, T v) const {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <class T>
    __device__ __forceinline__ T fmin(UA_GLOBAL T *p, T v) const {
        return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};
// LDS words of the ring's counters and round slots: ds_* instructions with 32-bit
// addresses, not flat ones (a flat access takes the vector memory path even when it
// resolves to LDS).
template <class T>
__device__ __forceinline__ T lds_ld(const UA_LDS T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void lds_st(UA_LDS T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T lds_add(UA_LDS T *p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The ring runs on one-block teams only (the planner gives a grid a ring only for a
// 16-wave team): counters and round slots in LDS, block barriers. The counters take the
// start of the block's LDS insert-buffer area (that path does not use the buffer), the
// continuation's parked successors follow them.
constexpr int kLdsSegCap = kBlockWaves * 64 * 8;  // keys of one round's LDS insert-buffer half: 16 waves x 64 pops x 8 moves
static_assert(kBucketMetaWords + GridProb::MAXS * kBlockWaves * 64 <= 2 * kLdsSegCap,
              "ring counters and parked successors fit the LDS buffer");
template <class Team>
__device__ __forceinline__ void team_solve_buckets(const QueryDesc &q, QueryResult *res, const Team &team,
                                                   const GridProb &P, UA_LDS BucketMeta *bm,
                                                   UA_LDS RoundSlot *slots) {
    const WgScope sc;
    const int l = lane_id();
    const int tw = team.tw, W = team.W;
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    uint64_t *const stats = team.m.stats;
    if (P.trivial()) {  // start == goal: cost 0, no search
        if (team.leader()) {
            res->status = ST_OK;
            res->found = 1;
            res->goal = 0;
            res->rounds = 0;
            res->pops = res->expanded = res->generated = res->inserted = res->atomics = 0;
            res->flags = 0;
        }
        return;
    }
    constexpr uint32_t RM = kBuckets - 1;
    const int bsh = q.cbits + q.bshift;  // key >> bsh = the key's bucket
    const int blog = q.bkt_log2;
    const uint32_t bmask = (1u << blog) - 1, bcap = 1u << blog;
    UA_GLOBAL uint64_t *const store = gptr(q.heap) + kBucketMetaWords;
    // per lane, MAXS successors parked in LDS after the counters while children expand
    UA_LDS uint64_t *const stash = reinterpret_cast<UA_LDS uint64_t *>(bm) + kBucketMetaWords;
    const bool resumed = q.resume != nullptr && q.resume[0] != 0;
    const int r0 = resumed ? (int)q.resume[1] : 0;
    // resume record after its header: the lowest bucket, then tail[] and the next round's
    // consumed counts as 32-bit words
    uint32_t *const rec32 = q.resume ? reinterpret_cast<uint32_t *>(q.resume + RESUME_HEADER + 1) : nullptr;
    uint64_t cur;  // the lowest bucket that can hold a key (wave-uniform)
    if (resumed) {
        cur = q.resume[RESUME_HEADER];
        for (int i = tid; i < kBuckets; i += nthr) {
            lds_st(&bm->tail[i], rec32[i]);
            lds_st(&bm->head[r0 & 1][i], rec32[kBuckets + i]);
        }
    } else {
        cur = P.item(octile_key(q.sx, q.sy, q.ex, q.ey), q.sx, q.sy) >> bsh;
        for (int i = tid; i < kBuckets; i += nthr) {
            lds_st(&bm->tail[i], 0u);
            lds_st(&bm->head[0][i], 0u);
        }
    }
    if (team.leader()) {
        for (int i = 0; i < 2; ++i) {
            lds_st(&slots[i].lb, SENT);
            lds_st(&slots[i].goal, SENT);
            lds_st(&slots[i].cnt, 0u);
            lds_st(&slots[i].err, 0u);
        }
        for (int i = 0; i < 4; ++i) team.sc.store(&stats[i], resumed ? q.resume[3 + i] : (uint64_t)0);
        team.sc.store(&stats[4], resumed ? q.resume[8] : (uint64_t)0);
    }
    __syncthreads();
    if (!resumed && team.leader()) {  // the start's record and key
        const uint64_t key0 = P.seed(team.sc);
        const uint32_t ring = (uint32_t)cur & RM;
        sc.store(store + ((uint64_t)ring << blog), key0);
        lds_st(&bm->tail[ring], 1u);
    }
    __syncthreads();

    uint32_t n_pop = 0;
    Ctr k;
    uint64_t best_goal = resumed ? q.resume[2] : SENT;
    uint32_t pruned = resumed ? (uint32_t)q.resume[7] & FLAG_PRUNED : 0u;
    int32_t status = ST_OK;
    int found = 0;
    const uint32_t pmax = (uint32_t)W * 64;  // keys a round takes at most
    const int win = q.bwin < 1 ? 1 : q.bwin;
    TeamVars *const tvp = team.tv;
    int r = r0;
    for (;; ++r) {
        UA_LDS RoundSlot *const cs = &slots[r & 1];
        UA_LDS uint32_t *const hcur = bm->head[r & 1];
        UA_LDS uint32_t *const hnext = bm->head[(r + 1) & 1];
        // (A) the ring's counters in window order: element i of lane l is bucket cur + 4 l + i
        uint32_t hd[4], lv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t ring = ((uint32_t)cur + 4u * (uint32_t)l + (uint32_t)i) & RM;
            hd[i] = lds_ld(&hcur[ring]);
            lv[i] = lds_ld(&bm->tail[ring]) - hd[i];
        }
        __syncthreads();
        // (B) plan
        int min_d = kBuckets;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t m = __ballot(lv[i] != 0);
            if (m) {
                const int d = 4 * __builtin_ctzll(m) + i;
                min_d = d < min_d ? d : min_d;
            }
        }
        if (min_d == kBuckets) {  // no live bucket: the open list is exhausted
            found = best_goal != SENT;
            status = found ? ST_OK : ST_NO_PATH;
            break;
        }
        const uint64_t blo = cur + (uint64_t)min_d;
        if (best_goal != SENT && best_goal <= ((blo << q.bshift) << q.qshift)) {  // goal <= every open f
            found = 1;
            break;
        }
        if (r + 1 >= q.max_rounds) {
            status = ST_ROUNDS;
            break;
        }
        if (team.leader()) {  // the next round's slot, and this round's time budget
            UA_LDS RoundSlot *const ns = &slots[(r + 1) & 1];
            lds_st(&ns->lb, SENT);
            lds_st(&ns->goal, SENT);
            lds_st(&ns->cnt, 0u);
            lds_st(&ns->err, 0u);
            const uint64_t dl = __hip_atomic_load(&tvp->deadline, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (dl && (uint64_t)now_ticks() > dl)
                __hip_atomic_fetch_or(&cs->err, (uint32_t)FLAG_TIMEUP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // the window's keys, clamped to pmax in window order
        uint32_t tk[4], lsum = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int d = 4 * l + i;
            tk[i] = (d >= min_d && d < min_d + win) ? lv[i] : 0u;
            lsum += tk[i];
        }
        int tot;
        const uint32_t ex = (uint32_t)wave_excl_scan((int)lsum, &tot);
        const uint32_t total = (uint32_t)tot < pmax ? (uint32_t)tot : pmax;
        uint32_t run = ex;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t lo = run < pmax ? run : pmax;
            run += tk[i];
            const uint32_t hi = run < pmax ? run : pmax;
            tk[i] = hi - lo;
        }
        const uint32_t exc = ex < pmax ? ex : pmax;  // this lane's first rank
        if (tw == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                lds_st(&hnext[((uint32_t)cur + 4u * (uint32_t)l + (uint32_t)i) & RM], hd[i] + tk[i]);
        }
        // (C) the key of rank j = 64 w + l: the lane L that holds it (the largest L with
        // exc_L <= j), then its element
        const uint32_t j = (uint32_t)tw * 64u + (uint32_t)l;
        int L = 0;
#pragma unroll
        for (int st = 32; st > 0; st >>= 1) {
            const int c = L + st;  // at most 63
            if ((uint32_t)__shfl((int)exc, c) <= j) L = c;
        }
        uint32_t off = j - (uint32_t)__shfl((int)exc, L);
        const uint32_t t0 = (uint32_t)__shfl((int)tk[0], L), t1 = (uint32_t)__shfl((int)tk[1], L),
                       t2 = (uint32_t)__shfl((int)tk[2], L);
        const uint32_t h0 = (uint32_t)__shfl((int)hd[0], L), h1 = (uint32_t)__shfl((int)hd[1], L),
                       h2 = (uint32_t)__shfl((int)hd[2], L), h3 = (uint32_t)__shfl((int)hd[3], L);
        int e = 0;
        uint32_t hsel = h0;
        if (off >= t0) {
            off -= t0;
            e = 1;
            hsel = h1;
            if (off >= t1) {
                off -= t1;
                e = 2;
                hsel = h2;
                if (off >= t2) {
                    off -= t2;
                    e = 3;
                    hsel = h3;
                }
            }
        }
        uint64_t it = SENT;
        if (j < total) {
            const uint32_t ring = ((uint32_t)cur + 4u * (uint32_t)L + (uint32_t)e) & RM;
            it = sc.load(store + (((uint64_t)ring << blog) | ((hsel + off) & bmask)));
        }
        cur = blo;
        // (D) expansion
        uint64_t s[GridProb::MAXS];
#pragma unroll
        for (int i = 0; i < GridProb::MAXS; ++i) s[i] = SENT;
        int c = 0;
        uint64_t goalmin = SENT;
        uint32_t err = 0;
        uint64_t gk = 0;
        if (it != SENT) {
            ++n_pop;
            if (P.dominated(it, best_goal)) it = SENT;
        }
        if (it != SENT) P.expand(sc, it, s, c, goalmin, err, k, &gk, best_goal);
        // continuation only in a saturated round (the window held more keys than the round
        // takes), so that the chains run ahead of a best-first frontier. While the children
        // are expanded, the other lanes' successors wait in LDS rather than in registers.
        uint64_t keep = SENT;  // the child's sibling, if any
        int dc = 0;
        const uint64_t child = it != SENT && (uint32_t)tot > pmax ? P.pick_child(it, s, c, &dc) : SENT;
        if (__ballot(child != SENT)) {
            if (child != SENT) {
#pragma unroll
                for (int i = 0; i < GridProb::MAXS; ++i) {
                    if (i != dc && s[i] < keep) keep = s[i];
                    s[i] = SENT;
                }
                c = keep != SENT;
                P.expand_child(sc, child, gk + (dc < 4 ? KEY_ONE : KEY_SQRT2), s, c, goalmin, err, k);
            } else {
#pragma unroll
                for (int i = 0; i < GridProb::MAXS; ++i) stash[i * nthr + tid] = s[i];
            }
            if (child == SENT) {
#pragma unroll
                for (int i = 0; i < GridProb::MAXS; ++i) s[i] = stash[i * nthr + tid];
            }
        }
        // (E) appends (s[] and the sibling kept by continuation): every key's tail add is
        // issued before any of the positions is used
        constexpr int NK = GridProb::MAXS + 1;
        uint32_t pos[NK];
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            uint64_t &key = i < GridProb::MAXS ? s[i] : keep;
            pos[i] = 0;
            if (key == SENT) continue;
            const uint64_t b = key >> bsh;
            if (b - cur >= (uint64_t)kBuckets) {  // outside the ring's window: impossible with a consistent h
                err |= FLAG_ABORT;
                key = SENT;
                continue;
            }
            pos[i] = lds_add(&bm->tail[(uint32_t)b & RM], 1u);
        }
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            const uint64_t key = i < GridProb::MAXS ? s[i] : keep;
            if (key == SENT) continue;
            const uint32_t ring = (uint32_t)(key >> bsh) & RM;
            // at least this many of the bucket's keys are unconsumed: full
            if (pos[i] - lds_ld(&hcur[ring]) >= bcap) {
                err |= FLAG_CAPACITY;
                continue;
            }
            sc.store(store + (((uint64_t)ring << blog) | (pos[i] & bmask)), key);
        }
        if (__ballot(goalmin != SENT)) {
            goalmin = wave_min64(goalmin);
            if (l == 0) __hip_atomic_fetch_min(&cs->goal, goalmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (__ballot(err != 0)) {
            for (int m = 32; m > 0; m >>= 1) err |= __shfl_xor(err, m);
            if (l == 0) __hip_atomic_fetch_or(&cs->err, err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // every wave reads the same slot values after the barrier
        __syncthreads();
        const uint64_t g = lds_ld(&cs->goal);
        const uint32_t rerr = lds_ld(&cs->err);
        best_goal = best_goal < g ? best_goal : g;
        pruned |= rerr & FLAG_PRUNED;
        if (rerr & ~(uint32_t)(FLAG_TIMEUP | FLAG_PRUNED)) {
            status = (rerr & FLAG_ABORT) ? ST_ABORT : (rerr & FLAG_CAPACITY) ? ST_CAPACITY : ST_RANGE;
            break;
        }
        if (q.resume != nullptr && ((rerr & FLAG_TIMEUP) || (q.round_budget > 0 && r + 1 - r0 >= q.round_budget))) {
            status = ST_SUSPENDED;  // every wave reaches this with the same r and slot values
            break;
        }
    }
    // ---- statistics, resume record and result
    const int64_t t_pop = wave_sum64(n_pop), t_exp = wave_sum64(k.exp), t_gen = wave_sum64(k.gen),
                  t_ins = wave_sum64(k.ins), t_atom = wave_sum64(k.atom);
    if (l == 0) {
        team.sc.fadd(&stats[0], (uint64_t)t_pop);
        team.sc.fadd(&stats[1], (uint64_t)t_exp);
        team.sc.fadd(&stats[2], (uint64_t)t_gen);
        team.sc.fadd(&stats[3], (uint64_t)t_ins);
        team.sc.fadd(&stats[4], (uint64_t)t_atom);
    }
    if (status == ST_SUSPENDED) {  // the ring after round r, with the next round's consumed counts
        for (int i = tid; i < kBuckets; i += nthr) {
            rec32[i] = lds_ld(&bm->tail[i]);
            rec32[kBuckets + i] = lds_ld(&bm->head[(r + 1) & 1][i]);
        }
    }
    __syncthreads();
    if (team.leader()) {
        if (status == ST_SUSPENDED) {
            q.resume[1] = (uint64_t)(r + 1);
            q.resume[2] = best_goal;
            for (int i = 0; i < 4; ++i) q.resume[3 + i] = team.sc.load(&stats[i]);
            q.resume[7] = pruned;
            q.resume[8] = team.sc.load(&stats[4]);
            q.resume[RESUME_HEADER] = cur;
            q.resume[0] = 1;
        }
        res->status = status;
        res->found = found;
        res->goal = best_goal;
        res->rounds = r + 1;
        res->pops = (int64_t)team.sc.load(&stats[0]);
        res->expanded = (int64_t)team.sc.load(&stats[1]);
        res->generated = (int64_t)team.sc.load(&stats[2]);
        res->inserted = (int64_t)team.sc.load(&stats[3]) + 1;
        res->atomics = (int64_t)team.sc.load(&stats[4]);
        res->flags = pruned;
    }
}

// A query of the other problem set (set_of) fails with ST_INVALID.
template <int SET, class Team>
__device__ __forceinline__ void dispatch(const QueryDesc &q, QueryResult *res, const Team &team) {
    if (set_of(q) != SET) {
        if (team.leader()) {
            res->status = ST_INVALID;
            res->found = 0;
        }
        return;
    }
    if constexpr (SET == SET_FIVE) {
        PuzzleProb5 P(q);
        team_solve(q, res, team, P);
    } else if constexpr (SET == SET_GHEAP) {
        GridProb P(q);
        team_solve(q, res, team, P);
    } else if (q.kind == KIND_GRID) {
        GridProb P(q);
        if (team.bpg != 1 || team.m.aux == nullptr) {  // a ring needs a one-block team with LDS counters
            if (team.leader()) {
                res->status = ST_INVALID;
                res->found = 0;
            }
            return;
        }
        team_solve_buckets(q, res, team, P, (UA_LDS BucketMeta *)team.m.aux, (UA_LDS RoundSlot *)team.m.slot);
    } else {
        PuzzleProb P(q, q.pdb);
        team_solve(q, res, team, P);
    }
}


// The query loop of one group: the group leader takes the next query index of its
// class (the hook runs on one thread inside the group barrier) and the group solves it
// as one team, until the class's queries are exhausted.
// With a time budget the group's first hook fixes its deadline; after the deadline it
// takes no more queries, and the round barrier of the query it runs flags FLAG_TIMEUP.
template <int SET, class Team>
__device__ __forceinline__ void group_loop(const QueryDesc *qs, QueryResult *rs, const Team &team, TeamVars *tv, uint32_t *head,
                           int nq, int qoff, int64_t budget) {
    for (;;) {
        int64_t t0 = now_ticks();
        if (!team.sync([tv, head, budget] {
                constexpr int SC = __HIP_MEMORY_SCOPE_AGENT;
                uint32_t qi = 0xFFFFFFFFu;  // none
                uint64_t now = (uint64_t)now_ticks();
                uint64_t dl = __hip_atomic_load(&tv->deadline, __ATOMIC_RELAXED, SC);
                if (budget > 0 && dl == 0) {
                    dl = now + (uint64_t)budget;
                    __hip_atomic_store(&tv->deadline, dl, __ATOMIC_RELAXED, SC);
                }
                if (dl == 0 || now <= dl) qi = atomicAdd(head, 1u);
                __hip_atomic_store(&tv->next_q, qi, __ATOMIC_RELAXED, SC);
            }))
            break;
        uint32_t k = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)__hip_atomic_load(&tv->next_q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (k >= (uint32_t)nq) break;
        int qi = qoff + k;
        QueryDesc q = qs[qi];
        dispatch<SET>(q, rs + qi, team);
        if (team.leader()) {
            rs[qi].t_start = t0;
            rs[qi].t_end = now_ticks();
        }
    }
}


// Body of k_teams for one block: find this block's class and group from the launch
// shape, then run the group's query loop.
template <int SET = SET_MAIN>
__device__ __forceinline__ void group_block(const QueryDesc *qs, QueryResult *rs, const LaunchShape &shape,
                                            uint32_t *qhead, TeamVars *tvs, uint64_t *segs) {
    int rel = (int)blockIdx.x, cls = 0, group = 0;
    if (shape.iso_xcds > 0) {
        // XCD split: the block's side (XCD blockIdx mod kXcds), its index among that side's
        // blocks (interleaved across the side's XCDs), then its class among the side's
        const int x = (int)blockIdx.x % kXcds, s = (int)blockIdx.x / kXcds;
        const int P = kXcds - shape.iso_xcds;
        const bool split_off = x >= P;
        rel = split_off ? s * shape.iso_xcds + (x - P) : s * P + x;
        cls = -1;
        for (int k = 0; k < kClasses; ++k) {
            if ((((shape.iso_mask >> k) & 1) != 0) != split_off) continue;
            const int n = shape.groups[k] * shape.bpg[k];
            if (rel < n) {
                cls = k;
                break;
            }
            rel -= n;
        }
        if (cls < 0) return;  // past its side's teams: idle
    } else {
        while (cls + 1 < kClasses && rel >= shape.groups[cls] * shape.bpg[cls]) {
            rel -= shape.groups[cls] * shape.bpg[cls];
            ++cls;
        }
    }
    size_t seg_off = 0;
    for (int k = 0; k < cls; ++k) {
        group += shape.groups[k];
        seg_off += (size_t)shape.groups[k] * 2 * (size_t)shape.segcap[k];
    }
    const int bpg = shape.bpg[cls];
    group += rel / bpg;
    const int segcap = shape.segcap[cls];
    seg_off += (size_t)(rel / bpg) * 2 * (size_t)segcap;
    // A one-block team keeps its round slots and insert buffer in LDS: the appends, the
    // slot atomics and the buffer reads of a round never leave the CU. (The accesses go
    // through generic pointers, so the same code serves LDS and global memory.)
    __shared__ RoundSlot view;
    __shared__ RoundSlot lds_slot[2];
    __shared__ uint64_t lds_seg[2 * kLdsSegCap];
    TeamVars *tv = tvs + group;
    TeamMem m;
    const bool in_lds = bpg == 1 && segcap <= kLdsSegCap;
    m.slot = in_lds ? lds_slot : tv->slot;
    m.view = &view;
    m.stats = &tv->n_pop;
    m.seg = in_lds ? lds_seg : segs + seg_off;
    m.segcap = segcap;
    // a one-block bucket team's counters take the LDS insert buffer's place (that path does
    // not use the buffer)
    static_assert(sizeof(BucketMeta) <= sizeof(lds_seg), "bucket counters fit the LDS buffer");
    m.aux = in_lds ? static_cast<void *>(lds_seg) : nullptr;
    GroupTeam team(m, tv, bpg, rel % bpg);
    group_loop<SET>(qs, rs, team, tv, qhead + cls, shape.nq[cls], shape.qoff[cls], shape.budget_ticks);
}

}  // namespace ua
