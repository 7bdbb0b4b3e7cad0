"""CPU run of the device search code (uastar_amd/csrc/search.hpp) under the SIMT
emulator in tests/emu/: one workgroup team, host memory, results against the oracle.

This checks the batch path's logic (one-block groups of k_teams, the group
barrier and query hand-off, per-wave heaps with HBM levels) without a GPU.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import goal_board, grid_bytes, random_grid

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
LIB = os.path.join(EMU, "libemu.so")
CSRC = os.path.join(os.path.dirname(HERE), "uastar_amd", "csrc")


class QueryDesc(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("kind", "n", "H", "W", "sx", "sy", "ex", "ey", "wpr", "ybits", "cbits",
                                          "qshift", "heap_cap", "hash_log2", "max_rounds", "heaps")] + [
        ("bitmap", C.c_void_p), ("best", C.c_void_p), ("vals", C.c_void_p), ("heap", C.c_void_p),
        ("start_state", C.c_uint64), ("start_state2", C.c_uint64), ("pdb", C.c_void_p), ("resume", C.c_void_p),
        ("round_budget", C.c_int32), ("cls", C.c_int32), ("wshift", C.c_int32), ("bounded", C.c_int32),
        ("open_kind", C.c_int32), ("bkt_log2", C.c_int32), ("bshift", C.c_int32), ("bwin", C.c_int32)]

BUCKETS, BUCKET_META_WORDS = 256, 512

RESUME_WORDS_4 = 9 + 4 * 130


class QueryResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("found", C.c_int32), ("goal", C.c_uint64), ("rounds", C.c_int64),
                ("pops", C.c_int64), ("expanded", C.c_int64), ("generated", C.c_int64), ("inserted", C.c_int64),
                ("atomics", C.c_int64), ("flags", C.c_uint32), ("pad", C.c_int32), ("t_start", C.c_int64), ("t_end", C.c_int64)]


@pytest.fixture(scope="module")
def emu():
    srcs = [os.path.join(EMU, "emu_batch.cpp"), os.path.join(EMU, "hip_emu.hpp")] + [
        os.path.join(CSRC, f) for f in ("search.hpp", "wave_heap.hpp", "common.hpp")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(s) for s in srcs):
        subprocess.check_call(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-pthread", "-I", CSRC,
                               srcs[0], "-o", LIB + ".tmp"])
        os.replace(LIB + ".tmp", LIB)
    L = C.CDLL(LIB)
    L.emu_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.emu_batch_gheap.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    return L


def bits_for(v):
    b = 1
    while (1 << b) < v:
        b += 1
    return b


class Batch:
    """Host-memory version of the runtime's batch planning (best tables, bucket rings or
    heaps). Grids use the team's bucket ring (kernel set SET_MAIN) unless grid_heaps, then
    one heap per wave (SET_GHEAP)."""

    def __init__(self, waves, max_rounds=100_000, grid_heaps=False, bwin=160):
        self.waves, self.max_rounds, self.grid_heaps, self.bwin = waves, max_rounds, grid_heaps, bwin
        self.descs, self.keep = [], []

    def add_grid(self, graph, H, W, sx, sy, ex, ey, items=65536, bkt_log2=None):
        bm = U.bytes_to_bitmap(graph, H, W)
        best = np.full(H * W, np.uint64(~np.uint64(0)))
        best[np.asarray(graph).reshape(-1) == 0] ^= np.uint64(16)  # each word carries its cell's open bit
        d = QueryDesc(kind=1, H=H, W=W, sx=sx, sy=sy, ex=ex, ey=ey, wpr=(W + 63) // 64, ybits=bits_for(W))
        d.cbits = bits_for(H) + d.ybits
        d.qshift = max(0, d.cbits - 10)
        d.max_rounds, d.heaps = self.max_rounds, self.waves
        if self.grid_heaps:
            per_heap = (items + 64 * self.waves - 1) // (64 * self.waves)
            cap = max(per_heap + 1, 4)
            heap = np.zeros(self.waves * (cap - 1) * 64, np.uint64)
            d.heap_cap = cap
        else:  # as grid_fill_desc: 1/16-step buckets, 1/8 of the capacity each (at least 4096)
            d.open_kind, d.bshift, d.bwin = 1, max(0, 32 - d.qshift), self.bwin
            d.bkt_log2 = bkt_log2 if bkt_log2 is not None else bits_for(max(4096, items // 8))
            heap = np.zeros(BUCKET_META_WORDS + (BUCKETS << d.bkt_log2), np.uint64)
        d.bitmap, d.best, d.heap = bm.ctypes.data, best.ctypes.data, heap.ctypes.data
        self.keep += [bm, best, heap]
        self.descs.append(d)

    def add_puzzle(self, n, tiles, pdb, hash_log2=17, items=65536):
        slots = np.zeros(2 << hash_log2, np.uint64)  # 16-byte slots: board, complemented value
        per_heap = (items + 64 * self.waves - 1) // (64 * self.waves)
        cap = max(per_heap + 1, 4)
        heap = np.zeros(self.waves * (cap - 1) * 64, np.uint64)
        s = 0
        for i, t in enumerate(tiles):
            s |= int(t) << (4 * i)
        d = QueryDesc(kind=2, n=n, hash_log2=hash_log2, heap_cap=cap, max_rounds=self.max_rounds, heaps=self.waves)
        d.best, d.heap, d.start_state = slots.ctypes.data, heap.ctypes.data, s
        d.pdb = pdb.ctypes.data
        self.keep += [slots, heap, pdb]
        self.descs.append(d)

    def run(self, lib):
        arr = (QueryDesc * len(self.descs))(*self.descs)
        res = (QueryResult * len(self.descs))()
        run = lib.emu_batch_gheap if self.grid_heaps else lib.emu_batch
        assert run(arr, res, len(self.descs), self.waves) == 0
        return list(res)


@pytest.mark.parametrize("waves,grid_heaps", [(4, False), (16, False), (4, True), (16, True)])
def test_emulated_batch_grids(emu, oracle, waves, grid_heaps):
    rng = np.random.default_rng(42)
    b = Batch(waves=waves, grid_heaps=grid_heaps)
    exp = []
    for i in range(6):
        H, W = int(rng.integers(2, 60)), int(rng.integers(2, 60))
        m = random_grid(rng, H, W, float(rng.uniform(0.0, 0.4)))
        sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
        m[sx, sy] = True
        g = grid_bytes(m)
        b.add_grid(g, H, W, sx, sy, ex, ey, items=4096)   # small heaps: HBM levels are used
        exp.append(oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref"))
    res = b.run(emu)
    for i, (r, e) in enumerate(zip(res, exp)):
        assert r.status in (0, 1), (i, r.status)
        assert bool(r.found) == e["found"], i
        if e["found"]:
            assert U.key_to_pair(r.goal) == (e["orth"], e["diag"]), i


@pytest.mark.parametrize("bwin", [1, 16, 160, 208])
def test_emulated_bucket_ring_grids(emu, oracle, bwin):
    """The bucket ring on 40 random grids (2..120 cells a side, obstacle rate 0..0.45,
    random endpoints) with narrow and wide round windows: costs equal the oracle's, and
    the expansions stay within bounds of the oracle A*'s (a narrow window pops near-best-
    first; a wide one takes whatever the round can hold). A query
    can exceed the oracle where many cells tie at the optimal f (a start whose octile
    distance is the optimum: every cell on any optimal path has f = C*, and the ring
    expands the whole tied bucket, the oracle's tie order only some of it)."""
    rng = np.random.default_rng(77 + bwin)
    b = Batch(waves=4, bwin=bwin)
    exp = []
    for i in range(40):
        H, W = int(rng.integers(2, 121)), int(rng.integers(2, 121))
        m = random_grid(rng, H, W, float(rng.uniform(0.0, 0.45)))
        sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
        m[sx, sy] = True
        g = grid_bytes(m)
        b.add_grid(g, H, W, sx, sy, ex, ey)
        exp.append(oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref"))
    res = b.run(emu)
    for i, (r, e) in enumerate(zip(res, exp)):
        assert r.status in (0, 1), (i, r.status)
        assert bool(r.found) == e["found"], i
        if e["found"]:
            assert U.key_to_pair(r.goal) == (e["orth"], e["diag"]), i
            assert r.expanded <= 24 * (e["expanded"] + 64), (i, r.expanded, e["expanded"])
    # a wide window on a small open list pops everything, as the heaps do (which expand
    # about 8x the oracle's on these grids)
    total, total_oracle = sum(r.expanded for r in res), sum(e["expanded"] for e in exp)
    assert total <= (4.0 if bwin > 16 else 1.5) * total_oracle, (total, total_oracle)


def test_emulated_bucket_ring_overflow_is_reported(emu, oracle):
    """A bucket too small for the search's keys fails the query with ST_CAPACITY (the
    runtime retries it larger) instead of overwriting unconsumed keys."""
    H = W = 90
    g = grid_bytes(np.ones((H, W), bool))
    b = Batch(waves=4)
    b.add_grid(g, H, W, 0, 45, 89, 45, bkt_log2=2)
    b.add_grid(g, H, W, 0, 45, 89, 45)
    res = b.run(emu)
    assert res[0].status == -3
    assert res[1].status == 0 and U.key_to_pair(res[1].goal) == (89, 0)


def test_emulated_batch_puzzles(emu, oracle):
    pdb3 = oracle.pdb_full(3)
    b = Batch(waves=4)
    exp = []
    for s in range(4):
        t = oracle.puzzle_random(3, 900 + s)
        b.add_puzzle(3, t, pdb3)
        exp.append(oracle.puzzle_solve(3, t, pdb3, "astar")[0])
    res = b.run(emu)
    for r, e in zip(res, exp):
        assert r.status == 0 and r.found and (r.goal >> 32) == e


def test_emulated_suspend_resume(emu, oracle):
    """Queries run a few rounds per launch, save their state to the resume record and
    continue in the next launch; the costs equal the oracle's."""
    rng = np.random.default_rng(9)
    b = Batch(waves=4)
    exp = []
    for i in range(3):
        H, W = int(rng.integers(20, 50)), int(rng.integers(20, 50))
        m = random_grid(rng, H, W, 0.25)
        m[0, 0] = True
        g = grid_bytes(m)
        b.add_grid(g, H, W, 0, 0, H - 1, W - 1, items=4096)
        exp.append(oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref"))
    pdb3 = oracle.pdb_full(3)
    t = oracle.puzzle_random(3, 31337)
    b.add_puzzle(3, t, pdb3)
    exp_p = oracle.puzzle_solve(3, t, pdb3, "astar")[0]
    recs = [np.zeros(RESUME_WORDS_4, np.uint64) for _ in b.descs]
    for d, r in zip(b.descs, recs):
        d.resume = r.ctypes.data
        d.round_budget = 3
    final = [None] * len(b.descs)
    launches = 0
    while any(f is None for f in final):
        idx = [i for i, f in enumerate(final) if f is None]
        sub = Batch(waves=4)
        sub.descs = [b.descs[i] for i in idx]
        res = sub.run(emu)
        launches += 1
        for i, r in zip(idx, res):
            if r.status != 2:  # not suspended
                final[i] = r
        assert launches < 200
    assert launches > 2
    for r, e in zip(final[:3], exp):
        assert r.status in (0, 1) and bool(r.found) == e["found"]
        if e["found"]:
            assert U.key_to_pair(r.goal) == (e["orth"], e["diag"])
    assert final[3].status == 0 and (final[3].goal >> 32) == exp_p


# ---------------------------------------------------------------- 5x5 boards

def pack5(tiles):
    """PuzzleStorage<5> layout (puzzle/storage.hpp:144-170): cells 0..11 in w0, 12..23 in
    w1, 5 bits each; cell 24's low 4 bits at w0 bit 60, its bit 4 at w1 bit 60."""
    t = [int(v) for v in tiles]
    w0 = sum(t[i] << (5 * i) for i in range(12)) | ((t[24] & 15) << 60)
    w1 = sum(t[i] << (5 * (i - 12)) for i in range(12, 24)) | ((t[24] & 16) << 56)
    return w0, w1


def short_walk(rng, length):
    b = list(range(1, 25)) + [0]
    z, last = 24, -1
    for _ in range(length):
        while True:
            d = int(rng.integers(4))
            x, y = z // 5 + (1, -1, 0, 0)[d], z % 5 + (0, 0, 1, -1)[d]
            if 0 <= x < 5 and 0 <= y < 5 and (last < 0 or d != (last ^ 1)):
                break
        nz = x * 5 + y
        b[z], b[nz] = b[nz], b[z]
        z, last = nz, d
    return b


@pytest.fixture(scope="module")
def emu5(emu):
    emu.emu_batch5.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    emu.emu_rank5.restype = C.c_uint32
    emu.emu_rank5.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    emu.emu_pos5.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    emu.emu_cell5.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    emu.emu_move5.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int, C.c_int]
    return emu


def test_board5_cells_positions_ranks(emu5, oracle):
    """The device's 5x5 board helpers against the oracle: every cell and tile position,
    the k-permutation rank of each 6-tile partition (database.cpp:140-163) and a blank
    move, on random boards and the goal."""
    parts = oracle.partitions(5)
    rng = np.random.default_rng(55)
    boards = [list(range(1, 25)) + [0]] + [oracle.puzzle_random(5, 7700 + i).tolist() for i in range(40)]
    for b in boards:
        w0, w1 = pack5(b)
        for c in range(25):
            assert emu5.emu_cell5(w0, w1, c) == b[c]
            assert emu5.emu_pos5(w0, w1, b[c]) == c
        for p, tracked in enumerate(parts):
            assert emu5.emu_rank5(w0, w1, p) == oracle.pdb_encode(5, tracked, np.array(b, np.uint8))
        z = b.index(0)
        for d in range(4):
            x, y = z // 5 + (1, -1, 0, 0)[d], z % 5 + (0, 0, 1, -1)[d]
            if not (0 <= x < 5 and 0 <= y < 5):
                continue
            nb = x * 5 + y
            a0, a1 = C.c_uint64(w0), C.c_uint64(w1)
            emu5.emu_move5(C.byref(a0), C.byref(a1), z, nb)
            c2 = list(b)
            c2[z], c2[nb] = c2[nb], c2[z]
            assert (a0.value, a1.value) == pack5(c2)
    assert rng is not None


def test_emulated_batch_puzzles5(emu5, oracle):
    """5x5 searches of the SET_FIVE kernel body on short walks off the goal, with the
    first 6-tile table of the reference's PDB (the other three zero: still admissible);
    steps equal IDA* with Manhattan distance."""
    parts = oracle.partitions(5)
    size = oracle.pdb_size(5, 6)
    pdb = np.zeros(4 * size, np.uint8)
    pdb[:size] = oracle.pdb_build(5, parts[0], 1)
    rng = np.random.default_rng(505)
    b = Batch(waves=4)
    exp = []
    for length in (0, 3, 8, 12, 14):
        t = short_walk(rng, length)
        keys = np.zeros(2 << 17, np.uint64)
        vals = np.full(1 << 17, 0xFFFFFFFF, np.uint32)
        cap = max((65536 + 64 * 4 - 1) // (64 * 4) + 1, 4)
        heap = np.zeros(4 * (cap - 1) * 64, np.uint64)
        d = QueryDesc(kind=2, n=5, hash_log2=17, heap_cap=cap, max_rounds=100_000, heaps=4)
        d.best, d.vals, d.heap, d.pdb = keys.ctypes.data, vals.ctypes.data, heap.ctypes.data, pdb.ctypes.data
        d.start_state, d.start_state2 = pack5(t)
        b.keep += [keys, vals, heap]
        b.descs.append(d)
        exp.append(oracle.puzzle_solve(5, np.array(t, np.uint8), None, "ida_manhattan")[0])
    b.keep.append(pdb)
    arr = (QueryDesc * len(b.descs))(*b.descs)
    res = (QueryResult * len(b.descs))()
    assert emu5.emu_batch5(arr, res, len(b.descs), 4) == 0
    for r, e in zip(res, exp):
        assert r.status == 0 and r.found and (r.goal >> 32) == e, (r.status, r.goal >> 32, e)
    # a 4x4 query sent to the 5x5 kernel fails cleanly
    b4 = Batch(waves=4)
    b4.add_puzzle(4, oracle.puzzle_random(4, 1), np.zeros(1, np.uint8))
    arr = (QueryDesc * 1)(*b4.descs)
    res = (QueryResult * 1)()
    assert emu5.emu_batch5(arr, res, 1, 4) == 0
    assert res[0].status == -1


def test_emulated_bounded_weighted_puzzles(emu, oracle):
    """The memory-bounded scheme's device side. Weighted keys (g >> wshift) + h with
    first-goal termination give solutions no shorter than the optimum, of the same
    parity. With bounded = 1 a hash too small for the search drops successors instead
    of failing: the query ends without an error, FLAG_PRUNED set; without bounded it
    fails with ST_CAPACITY."""
    pdb3 = oracle.pdb_full(3)
    b = Batch(waves=4)
    exp = []
    for s in range(3):
        t = oracle.puzzle_random(3, 4400 + s)
        exp.append(oracle.puzzle_solve(3, t, pdb3, "astar")[0])
        b.add_puzzle(3, t, pdb3, hash_log2=15)
        b.descs[-1].wshift = 1 + s
    for s, (ws, bounded) in enumerate([(1, 1), (0, 1), (0, 0)]):
        b.add_puzzle(3, oracle.puzzle_random(3, 4400 + s), pdb3, hash_log2=9)
        b.descs[-1].wshift, b.descs[-1].bounded = ws, bounded
    res = b.run(emu)
    for r, e in zip(res[:3], exp):
        assert r.status == 0 and r.found and r.flags == 0, r.status
        steps = r.goal >> 32
        assert steps >= e and (steps - e) % 2 == 0, (steps, e)
    for r in res[3:5]:  # bounded: pruned, no error
        assert r.status in (0, 1) and r.flags & 16
    assert res[5].status == -3  # unbounded: capacity error


@pytest.mark.slow
def test_emulated_fifteen_puzzles(emu, oracle):
    """15-puzzles, whose heuristic sums two PDB partitions: a child's h is its parent's
    with the moved tile's partition term swapped, the parent's second term read from its
    item key. The four fixture boards with the fewest expansions give their stored move
    counts, also when suspended every few rounds; a weighted run (wshift 1) stops at a
    solution no shorter than the optimum and of its parity."""
    pdb4 = oracle.pdb_full(4, 1, cache_dir="_cache")
    with open(os.path.join(HERE, "golden", "c3_puzzles.json")) as f:
        rows = sorted(json.load(f), key=lambda r: r["expanded"])[:4]
    b = Batch(waves=4)
    for r in rows:
        b.add_puzzle(4, np.array(r["board"], np.uint8), pdb4, hash_log2=18)
    b.add_puzzle(4, np.array(rows[0]["board"], np.uint8), pdb4, hash_log2=18)
    b.descs[-1].wshift = 1
    res = b.run(emu)
    for r, row in zip(res[:4], rows):
        assert r.status == 0 and r.found and (r.goal >> 32) == row["steps"], (row["seed"], r.status)
    steps = res[4].goal >> 32
    assert res[4].status == 0 and steps >= rows[0]["steps"] and (steps - rows[0]["steps"]) % 2 == 0
    # the same boards suspended every 5 rounds: the resumed keys carry their terms
    b2 = Batch(waves=4)
    for r in rows[:2]:
        b2.add_puzzle(4, np.array(r["board"], np.uint8), pdb4, hash_log2=18)
    recs = [np.zeros(RESUME_WORDS_4, np.uint64) for _ in b2.descs]
    for d, rec in zip(b2.descs, recs):
        d.resume, d.round_budget = rec.ctypes.data, 5
    final = [None] * len(b2.descs)
    for _ in range(400):
        idx = [i for i, f in enumerate(final) if f is None]
        if not idx:
            break
        sub = Batch(waves=4)
        sub.descs = [b2.descs[i] for i in idx]
        for i, r in zip(idx, sub.run(emu)):
            if r.status != 2:
                final[i] = r
    for r, row in zip(final, rows[:2]):
        assert r is not None and r.status == 0 and (r.goal >> 32) == row["steps"], row["seed"]


def test_emulated_heap_selftest(emu):
    """The per-wave batched heap on one emulated wave: random insert batches (0..64 keys,
    narrow and wide key ranges, so ties are common) and pops; every pop returns the 64
    smallest keys present, in order."""
    emu.emu_heap_selftest.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                      C.POINTER(C.c_int)]
    rng = np.random.default_rng(17)
    for trial in range(2):
        ops, keys, ref, pops = [], [], [], []
        for _ in range(110):
            if rng.random() < 0.7:
                c = int(rng.integers(0, 65))
                ks = rng.integers(0, [10, 1 << 62][trial], size=c, dtype=np.uint64)
                ops.append(c)
                keys.extend(ks.tolist())
                ref.extend(ks.tolist())
            else:
                ops.append(-1)
                ref.sort()
                pops.append(ref[:64])
                del ref[:64]
        ops_a = np.array(ops, np.int32)
        keys_a = np.array(keys, np.uint64)
        storage = np.zeros(4096 * 64, np.uint64)
        out = np.zeros(max(1, len(pops)) * 64, np.uint64)
        status = C.c_int()
        assert emu.emu_heap_selftest(ops_a.ctypes.data, len(ops), keys_a.ctypes.data, storage.ctypes.data, 4096,
                                     out.ctypes.data, C.byref(status)) == 0
        assert status.value == len(ref), trial
        for i, exp in enumerate(pops):
            row = [int(v) for v in out[i * 64:(i + 1) * 64] if int(v) != (1 << 64) - 1]
            assert row == [int(v) for v in exp], f"trial {trial} pop {i}"
