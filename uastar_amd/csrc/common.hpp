// common.hpp -- definitions shared by the HIP kernels and the host runtime.
//
// Exact grid costs. A path with `a` orthogonal and `b` diagonal steps costs
// a + b*sqrt(2) (utils.hpp:119-122 of uAstar uses FP32 for this). Here a cost is
// the integer key
//      key(a, b) = a * 2^36 + b * S,   S = 97184015999 = floor(sqrt(2) * 2^36) (odd)
// which is linear, so path keys add up exactly, step by step, in any order.
// For two costs with values below COST_LIMIT = 180000 the key order equals the
// exact order of a + b*sqrt(2): distinct values differ by at least 1/(2F), the
// key error is below b * 2^-36 <= F * 2^-36, and 2F^2 < 2^36 for F <= 185363.
// Because S is odd, (a, b) is recovered from the key modulo 2^36.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define UA_HD __host__ __device__
#else
#define UA_HD
#endif

namespace ua {

constexpr int WAVE = 64;
constexpr uint64_t KEY_ONE = 1ull << 36;
constexpr uint64_t KEY_SQRT2 = 97184015999ull;
constexpr uint64_t KEY_MASK36 = (1ull << 36) - 1;
constexpr double COST_LIMIT = 180000.0;
constexpr uint64_t KEY_LIMIT = 180000ull << 36;   // every f key must stay below this
constexpr uint64_t SENT = ~0ull;                  // empty heap slot / "infinity"
constexpr int kPuzzleSlotBits = 40;               // hash slot field of a puzzle item key

// Move table of uAstar (utils.hpp:120-121): 0..3 orthogonal, 4..7 diagonal.
// The same order serves the puzzle's blank moves (indices 0..3).
__constant__ static const int kDX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
__constant__ static const int kDY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

UA_HD inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// Grid generator spec: cell id (x*W+y) is blocked iff (cell_hash(seed,id) >> 32) < thr,
// thr = floor(block_rate * 2^32); start and goal are forced open.
UA_HD inline uint64_t cell_hash(uint64_t seed_mixed, uint64_t id) {
    return mix64(seed_mixed + id * 0x9E3779B97F4A7C15ull);
}

UA_HD inline uint64_t cost_key(uint64_t a, uint64_t b) { return a * KEY_ONE + b * KEY_SQRT2; }
UA_HD inline uint64_t octile_key(int x, int y, int ex, int ey) {
    int dx = x > ex ? x - ex : ex - x;
    int dy = y > ey ? y - ey : ey - y;
    int mn = dx < dy ? dx : dy;
    int df = dx > dy ? dx - dy : dy - dx;
    return cost_key((uint64_t)df, (uint64_t)mn);
}
// inverse of S modulo 2^36 (S odd), by Newton iteration
UA_HD inline uint64_t inv_sqrt2_mod36() {
    uint64_t x = KEY_SQRT2;  // correct to 3 bits
    for (int i = 0; i < 6; ++i) x *= 2 - KEY_SQRT2 * x;
    return x & KEY_MASK36;
}
UA_HD inline void key_to_pair(uint64_t key, int64_t *a, int64_t *b) {
    uint64_t bb = ((key & KEY_MASK36) * inv_sqrt2_mod36()) & KEY_MASK36;
    *b = (int64_t)bb;
    *a = (int64_t)((key - bb * KEY_SQRT2) >> 36);
}

// Grid best word, one per cell: (g_key << kGridKeyShift) | (open << 4) | dir, where dir
// is the move that reached the cell (8 = start, 15 = unseen). All ones with the open bit
// as the cell's is "unseen". g_key < KEY_LIMIT < 2^54, so the word never overflows.
constexpr int kGridKeyShift = 5;
constexpr uint64_t kOpenBit = 16;

// ------------------------------------------------------------------ query records
enum : int32_t { KIND_GRID = 1, KIND_PUZZLE = 2 };

// Status codes, identical to include/uastar.h
enum : int32_t {
    ST_OK = 0,
    ST_NO_PATH = 1,
    ST_INVALID = -1,
    ST_OOM = -2,
    ST_CAPACITY = -3,
    ST_RANGE = -4,
    ST_HIP = -5,
    ST_NO_DEVICE = -6,
    ST_ROUNDS = -7,
    ST_NO_PDB = -8,
    ST_ABORT = -9,
    ST_SUSPENDED = 2,   // round budget of a serving launch used up; state saved for resumption
};
// error bits raised by the kernels
// FLAG_TIMEUP is no error: the launch's time budget ran out (serving suspends the query).
// FLAG_PRUNED is no error either: a memory-bounded search dropped successors that did not
// fit its hash or heaps, so its answer is not proven optimal.
enum : uint32_t { FLAG_CAPACITY = 1, FLAG_RANGE = 2, FLAG_ABORT = 4, FLAG_TIMEUP = 8, FLAG_PRUNED = 16 };

// One search instance as the kernels see it. All pointers are device pointers.
struct QueryDesc {
    int32_t kind;
    int32_t n;               // puzzle side
    int32_t H, W;            // grid
    int32_t sx, sy, ex, ey;
    int32_t wpr;             // bitmap words per row
    int32_t ybits, cbits, qshift;  // grid item encoding
    int32_t heap_cap;        // full heap nodes per heap (root included)
    int32_t hash_log2;       // puzzle hash table size
    int32_t max_rounds;
    int32_t heaps;           // heaps (= waves) in the team; filled by the launcher
    const uint64_t *bitmap;  // grid: row-padded, 1 = open
    uint64_t *best;          // grid: per cell (gkey << 5 | open << 4 | dir);  puzzle: hash keys
    uint32_t *vals;          // puzzle: hash values (g << 4 | dir)
    uint64_t *heap;          // heap node storage, heaps * (heap_cap-1) * 64 u64
    uint64_t start_state;    // puzzle board (n = 5: its first word)
    uint64_t start_state2;   // n = 5: the board's second word
    const uint8_t *pdb;      // puzzle: the pattern databases of side n
    uint64_t *resume;        // serving: resume record (resume_words(heaps) u64), or null
    int32_t round_budget;    // serving: at most this many rounds per launch (0 = no limit)
    int32_t cls;             // host side: instance class of the launch (CLS_*)
    // puzzles, memory-bounded scheme: keys order by (g >> wshift) + h, i.e. weight
    // 2^wshift on h, and the search stops at its first goal when wshift > 0; bounded:
    // successors that do not fit the hash or the heaps are dropped (FLAG_PRUNED)
    int32_t wshift;
    int32_t bounded;
    // grids: the open list. OPEN_HEAPS = one batched heap per wave (heap, heap_cap);
    // OPEN_BUCKETS = a ring of kBuckets buckets shared by a one-block team, bucket
    // (f_key >> (qshift + bshift)) mod kBuckets, 2^bkt_log2 keys each, stored in `heap`;
    // a round takes keys from the lowest bwin buckets
    int32_t open_kind;
    int32_t bkt_log2;
    int32_t bshift;
    int32_t bwin;
};
enum : int32_t { OPEN_HEAPS = 0, OPEN_BUCKETS = 1 };
constexpr int kBuckets = 256;   // buckets of the ring (a power of two)
constexpr int kBucketMetaWords = 512;  // u64 words before a ring's keys, for its counters (4 KiB)

// One path walk of a finished query (k_paths): grid cell ids or puzzle blank moves,
// goal..start, at most len entries at offset off of the output; got = the count walked,
// -1 on a broken chain.
struct PathJob {
    const uint64_t *best;    // grid best table / puzzle hash keys
    const uint32_t *vals;    // puzzle hash values
    int32_t kind, n, H, W, sx, sy, ex, ey, hash_log2, pad;
    uint64_t goal_slot;      // puzzle: the goal's hash slot
    int64_t off, len, got;
};

// Problem sets of the search kernel: SET_MAIN solves grids on bucket rings and 3x3 / 4x4
// puzzles, SET_FIVE 5x5 puzzles, SET_GHEAP grids on per-wave heaps. Each set is its own
// kernel, so that no search raises the register demand of another (one kernel holding all
// of them spills heavily).
constexpr int SET_MAIN = 0, SET_FIVE = 1, SET_GHEAP = 2;
constexpr int kSets = 3;
UA_HD inline int set_of(const QueryDesc &q) {
    if (q.kind == KIND_PUZZLE) return q.n == 5 ? SET_FIVE : SET_MAIN;
    return q.open_kind == OPEN_HEAPS ? SET_GHEAP : SET_MAIN;
}

// The search kernel runs blocks of 16 wavefronts (1024 threads). A team is a group of
// whole blocks: 16, 32, 64 ... waves (= heaps).
constexpr int kBlockWaves = 16;

// Instance classes of a batched / served launch, each with its own team width. CLS_WIDE
// holds served queries that ran long and were widened to a large team.
constexpr int kClasses = 4;
enum : int { CLS_PUZZLE = 0, CLS_GRID = 1, CLS_GRID_LARGE = 2, CLS_WIDE = 3 };

// Shape of one k_teams launch. Class c has groups[c] groups of bpg[c] blocks (classes
// in order of c in block order); its groups pull the queries qs[qoff[c] .. qoff[c] +
// nq[c]) from counter qhead[c]. A group's insert buffer holds segcap[c] keys per
// round half.
//
// XCD split (iso_xcds > 0): the device's kXcds XCDs each have their own L2, and a
// workgroup b of a launch runs on XCD b mod kXcds. The classes of iso_mask (one-block
// teams, whose barriers never write back or invalidate an L2) then run on the last
// iso_xcds XCDs and every other class (teams of several blocks, whose agent-scope
// barriers do, and which stream pattern databases and hashes through the L2) on the
// first kXcds - iso_xcds; each side's classes fill its XCDs in class order, interleaved
// across them. Class order still fixes every group's index and insert buffer.
constexpr int kXcds = 8;
struct LaunchShape {
    int32_t groups[kClasses];
    int32_t bpg[kClasses];
    int32_t nq[kClasses];
    int32_t qoff[kClasses];
    int32_t segcap[kClasses];
    int16_t iso_mask;  // bit c: class c runs on the split-off XCDs
    int16_t iso_xcds;  // XCDs split off (0 = every class on every XCD)
    // serving: each group stops starting queries, and suspends the one it runs, once it
    // has been busy this long in the launch (s_memrealtime ticks, 100 MHz; 0 = no limit)
    int64_t budget_ticks;
};
UA_HD inline int shape_blocks(const LaunchShape &s) {
    int b = 0;
    for (int c = 0; c < kClasses; ++c) b += s.groups[c] * s.bpg[c];
    return b;
}

// The I/O block of one k_teams launch, in device memory and copied there from the host
// in one transfer before the launch: this header, the groups' TeamVars (zero), the
// launch's query descriptors and their results. A query the launch does not reach
// keeps the result the host put there.
struct StepHeader {
    LaunchShape shape;
    uint32_t qhead[kClasses];             // next query of each class (zero at launch)
    uint32_t tv_off, qd_off, qr_off, nq;  // byte offsets of TeamVars[], QueryDesc[], QueryResult[]
};

// Resume record of a suspended query: header, then per heap its root batch,
// partial batch and (n, pc). Written when a serving launch's round budget runs out.
constexpr int RESUME_HEADER = 9;   // valid, round, best_goal, pops, expanded, generated, inserted, flags, atomics
constexpr int RESUME_PER_HEAP = 130;
UA_HD inline int64_t resume_words(int heaps) { return RESUME_HEADER + (int64_t)heaps * RESUME_PER_HEAP; }

struct QueryResult {
    int32_t status;
    int32_t found;
    uint64_t goal;       // grid: goal g key; puzzle: (g << 32) | goal slot
    int64_t rounds;
    int64_t pops, expanded, generated, inserted, atomics;
    uint32_t flags;
    int32_t pad;
    int64_t t_start, t_end;   // device clock (s_memrealtime, 100 MHz) for latency
};

// Per-round values of a team. Round r uses slot r & 1; the barrier that ends
// round r resets slot (r + 1) & 1, so every wave reads identical values.
struct RoundSlot {
    uint64_t lb;     // min popped item key of the round (SENT = every heap was empty)
    uint64_t goal;   // min goal key found in the round
    uint32_t cnt;    // items appended to the insert buffer
    uint32_t err;    // FLAG_* raised between the previous barrier and this one
};

// Per-team scratch that lives across rounds (global memory).
struct TeamVars {
    RoundSlot slot[2];
    uint32_t bar_count;
    uint32_t bar_gen;
    uint32_t abort;
    uint32_t next_q;     // query index handed from the group leader to the group
    uint64_t n_pop, n_exp, n_gen, n_ins, n_atom;
    uint64_t deadline;   // serving: device time at which the group's budget ends (0 = unset)
};

}  // namespace ua
