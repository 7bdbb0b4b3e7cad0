"""uastar_amd -- Python binding of libuastar (the MI355X GA* engine, include/uastar.h).

The library is the product; this module only marshals arguments over ctypes.
It raises if libuastar.so is missing: there is no CPU fallback.

Classes mirror the reference's solver interface:
  Context.grid_solve / grid_path     <- GPUPathwaySolver::solve / getSolution
  Context.puzzle_solve / puzzle_moves <- GPUPuzzleSolver<N>::solve / getSolution
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuastar.so")

UA_OK = 0
UA_NO_PATH = 1
STATUS_NAMES = {0: "OK", 1: "NO_PATH", -1: "INVALID", -2: "OOM", -3: "CAPACITY", -4: "RANGE",
                -5: "HIP", -6: "NO_DEVICE", -7: "ROUNDS", -8: "NO_PDB", -9: "ABORT"}
QUERY_GRID = 1
PDB_BYTES = {3: 362880, 4: 576576000, 5: 510048000}  # the partitions' tables of side n, concatenated
QUERY_PUZZLE = 2
SQRT2 = math.sqrt(2.0)


class UAError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("team_waves", C.c_int32), ("plain_heaps", C.c_int32), ("max_rounds", C.c_int32),
                ("hash_log2", C.c_int32), ("mem_budget", C.c_uint64), ("puzzle_team_waves", C.c_int32),
                ("flags", C.c_int32)]


FLAG_GRAPH = 1
FLAG_SERVE_PATHS = 2
FLAG_GRID_HEAPS = 4
FLAG_XCD_SHARED = 8


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("found", C.c_int32), ("orth", C.c_int64), ("diag", C.c_int64),
                ("steps", C.c_int64), ("cost", C.c_double), ("rounds", C.c_int64), ("pops", C.c_int64),
                ("expanded", C.c_int64), ("generated", C.c_int64), ("inserted", C.c_int64),
                ("device_ms", C.c_double), ("kernel_ms", C.c_double), ("latency_ms", C.c_double),
                ("optimal", C.c_int32), ("weight", C.c_int32), ("atomics", C.c_int64), ("heaps", C.c_int32),
                ("reserved", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Query(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("grid", C.c_void_p), ("sx", C.c_int32),
                ("sy", C.c_int32), ("ex", C.c_int32), ("ey", C.c_int32), ("tiles", C.c_uint8 * 25)]


class Tuning(C.Structure):
    _fields_ = [("grid_team_waves", C.c_int32), ("grid_large_team_waves", C.c_int32),
                ("grid_large_cells", C.c_int64), ("puzzle_team_waves", C.c_int32), ("puzzle_hash_log2", C.c_int32),
                ("plain_grid_heaps", C.c_int32), ("plain_puzzle_heaps", C.c_int32), ("plain_hash_log2", C.c_int32),
                ("trials", C.c_int32), ("est_grid_ms", C.c_double), ("est_grid_large_ms", C.c_double),
                ("est_puzzle_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """Loads libuastar.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m uastar_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
        pp = C.POINTER(C.c_void_p)
        sig = {
            "ua_device_count": (C.c_int, []),
            "ua_create": (C.c_int, [C.c_int, C.POINTER(Config), pp]),
            "ua_destroy": (None, [vp]),
            "ua_last_error": (C.c_char_p, [vp]),
            "ua_grid_create": (C.c_int, [vp, C.c_int, C.c_int, vp, pp]),
            "ua_grid_from_bytes": (C.c_int, [vp, C.c_int, C.c_int, vp, pp]),
            "ua_grid_generate": (C.c_int, [vp, C.c_int, C.c_int, u64, C.c_double, C.c_int, C.c_int,
                                           C.c_int, C.c_int, pp]),
            "ua_grid_download": (C.c_int, [vp, vp, vp]),
            "ua_grid_free": (None, [vp, vp]),
            "ua_grid_solve": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Result)]),
            "ua_grid_path": (C.c_int, [vp, vp, i64, C.POINTER(i64)]),
            "ua_grid_solve_noalloc": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Result), vp,
                                                i64, C.POINTER(i64)]),
            "ua_puzzle_build_pdb": (C.c_int, [vp, C.c_int]),
            "ua_puzzle_download_pdb": (C.c_int, [vp, C.c_int, vp, u64]),
            "ua_puzzle_load_pdb": (C.c_int, [vp, C.c_int, vp, u64]),
            "ua_puzzle_solve": (C.c_int, [vp, C.c_int, vp, C.POINTER(Result)]),
            "ua_puzzle_moves": (C.c_int, [vp, vp, i64, C.POINTER(i64)]),
            "ua_puzzle_solve_bounded": (C.c_int, [vp, C.c_int, vp, u64, C.POINTER(Result)]),
            "ua_batch_solve": (C.c_int, [vp, C.POINTER(Query), i64, C.POINTER(Result)]),
            "ua_batch_path": (C.c_int, [vp, i64, vp, i64, C.POINTER(i64)]),
            "ua_serve_submit": (C.c_int, [vp, C.POINTER(Query), C.POINTER(u64)]),
            "ua_serve_step": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64)]),
            "ua_serve_poll": (C.c_int, [vp, C.POINTER(u64), C.POINTER(Result), i64, C.POINTER(i64)]),
            "ua_serve_drain": (C.c_int, [vp]),
            "ua_serve_path": (C.c_int, [vp, u64, vp, i64, C.POINTER(i64)]),
            "ua_serve_stats": (C.c_int, [vp, C.POINTER(i64), C.POINTER(C.c_double), C.POINTER(i32),
                                         C.POINTER(i64)]),
            "ua_launch_stats": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]),
            "ua_launch_graph_error": (C.c_char_p, [vp]),
            "ua_debug_atomic_rate": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "ua_debug_access_latency": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "ua_debug_heap_selftest": (C.c_int, [vp, vp, C.c_int, vp, i64, C.c_int, vp, i64,
                                                 C.POINTER(i32)]),
            "ua_autotune": (C.c_int, [vp, C.POINTER(Query), i64, C.POINTER(Tuning)]),
            "ua_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
            "ua_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    return lib().ua_device_count()


# ---------------------------------------------------------------- input helpers

def words_per_row(W: int) -> int:
    return (W + 63) // 64


def bytes_to_bitmap(graph: np.ndarray, H: int, W: int) -> np.ndarray:
    """Byte-per-cell graph (0 = blocked) -> row-padded 1-bit bitmap (uint64 words)."""
    g = (np.asarray(graph).reshape(H, W) != 0)
    wpr = words_per_row(W)
    pad = np.zeros((H, wpr * 64), dtype=bool)
    pad[:, :W] = g
    return np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(H, wpr).copy()


def bitmap_to_bytes(bm: np.ndarray, H: int, W: int) -> np.ndarray:
    bits = np.unpackbits(np.ascontiguousarray(bm).reshape(H, -1).view(np.uint8), axis=1, bitorder="little")
    return np.where(bits[:, :W] != 0, 0xFF, 0).astype(np.uint8).reshape(-1)


def mix64(z: int) -> int:
    M = (1 << 64) - 1
    z &= M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def cell_open(seed: int, block_rate: float, W: int, x: int, y: int) -> bool:
    """Host evaluation of the generator spec for one cell (used to pick open starts/goals)."""
    thr = int(math.floor(block_rate * 4294967296.0))
    h = mix64((mix64(seed) + (x * W + y) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1))
    return (h >> 32) >= thr


def key_to_pair(key: int):
    S = 97184015999
    inv = pow(S, -1, 1 << 36)
    b = ((key & ((1 << 36) - 1)) * inv) & ((1 << 36) - 1)
    a = (key - b * S) >> 36
    return a, b


@dataclass
class Grid:
    ctx: "Context"
    handle: C.c_void_p
    H: int
    W: int

    def download(self) -> np.ndarray:
        bm = np.empty((self.H, words_per_row(self.W)), np.uint64)
        self.ctx._check(lib().ua_grid_download(self.ctx.h, self.handle, bm.ctypes.data))
        return bm

    def to_bytes(self) -> np.ndarray:
        return bitmap_to_bytes(self.download(), self.H, self.W)

    def free(self):
        if self.handle:
            lib().ua_grid_free(self.ctx.h, self.handle)
            self.handle = None


class Context:
    """One device context (ua_ctx)."""

    def __init__(self, ordinal=0, team_waves=0, plain_heaps=0, max_rounds=0, hash_log2=0, mem_budget=0,
                 puzzle_team_waves=0, graphs=False, serve_paths=False, grid_heaps=False, xcd_shared=False):
        flags = ((FLAG_GRAPH if graphs else 0) | (FLAG_SERVE_PATHS if serve_paths else 0) |
                 (FLAG_GRID_HEAPS if grid_heaps else 0) | (FLAG_XCD_SHARED if xcd_shared else 0))
        self.cfg = Config(team_waves, plain_heaps, max_rounds, hash_log2, mem_budget, puzzle_team_waves, flags)
        h = C.c_void_p()
        rc = lib().ua_create(ordinal, C.byref(self.cfg), C.byref(h))
        if rc:
            raise UAError(rc, "ua_create failed")
        self.h = h

    def close(self):
        if self.h:
            lib().ua_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, ok=(UA_OK,)):
        if rc not in ok:
            raise UAError(rc, lib().ua_last_error(self.h).decode(errors="replace"))
        return rc

    # ---- grids
    def grid_from_bitmap(self, bm: np.ndarray, H: int, W: int) -> Grid:
        bm = np.ascontiguousarray(bm, dtype=np.uint64)
        assert bm.size == H * words_per_row(W)
        out = C.c_void_p()
        self._check(lib().ua_grid_create(self.h, H, W, bm.ctypes.data, C.byref(out)))
        return Grid(self, out, H, W)

    def grid_from_bytes(self, graph: np.ndarray, H: int, W: int) -> Grid:
        g = np.ascontiguousarray(graph, dtype=np.uint8).reshape(-1)
        assert g.size == H * W
        out = C.c_void_p()
        self._check(lib().ua_grid_from_bytes(self.h, H, W, g.ctypes.data, C.byref(out)))
        return Grid(self, out, H, W)

    def grid_generate(self, H, W, seed, block_rate, sx, sy, ex, ey) -> Grid:
        out = C.c_void_p()
        self._check(lib().ua_grid_generate(self.h, H, W, seed, block_rate, sx, sy, ex, ey, C.byref(out)))
        return Grid(self, out, H, W)

    def grid_solve(self, grid: Grid, sx, sy, ex, ey) -> dict:
        r = Result()
        self._check(lib().ua_grid_solve(self.h, grid.handle, sx, sy, ex, ey, C.byref(r)), (UA_OK, UA_NO_PATH))
        return r.as_dict()

    def grid_solve_noalloc(self, grid: Grid, sx, sy, ex, ey, want_path=True):
        """ua_grid_solve_noalloc: the plain solve and its path with no device allocation
        (grids of at most 2^20 cells). Returns (result dict, path or None)."""
        r = Result()
        n = C.c_int64()
        cap = grid.H * grid.W if want_path else 0
        ids = np.empty(max(1, cap), np.int32)
        rc = lib().ua_grid_solve_noalloc(self.h, grid.handle, sx, sy, ex, ey, C.byref(r),
                                         ids.ctypes.data if want_path else None, cap, C.byref(n))
        self._check(rc, (UA_OK, UA_NO_PATH))
        return r.as_dict(), (ids[: n.value].copy() if want_path and r.found else None)

    def grid_path(self) -> np.ndarray:
        n = C.c_int64()
        self._check(lib().ua_grid_path(self.h, None, 0, C.byref(n)))
        ids = np.empty(n.value, np.int32)
        self._check(lib().ua_grid_path(self.h, ids.ctypes.data, n.value, C.byref(n)))
        return ids

    # ---- puzzles
    def puzzle_build_pdb(self, n: int):
        self._check(lib().ua_puzzle_build_pdb(self.h, n))

    def puzzle_download_pdb(self, n: int) -> np.ndarray:
        size = PDB_BYTES[n]
        out = np.empty(size, np.uint8)
        self._check(lib().ua_puzzle_download_pdb(self.h, n, out.ctypes.data, size))
        return out

    def puzzle_load_pdb(self, n: int, pdb: np.ndarray):
        pdb = np.ascontiguousarray(pdb, dtype=np.uint8)
        self._check(lib().ua_puzzle_load_pdb(self.h, n, pdb.ctypes.data, pdb.size))

    def puzzle_solve(self, n: int, tiles) -> dict:
        t = np.ascontiguousarray(tiles, dtype=np.uint8).reshape(-1)
        r = Result()
        self._check(lib().ua_puzzle_solve(self.h, n, t.ctypes.data, C.byref(r)), (UA_OK, UA_NO_PATH))
        return r.as_dict()

    def puzzle_solve_bounded(self, n: int, tiles, mem_bytes: int) -> dict:
        """The memory-bounded scheme: plain A* in mem_bytes, else weighted searches that
        drop what does not fit; result["optimal"] says whether the cost is proven."""
        t = np.ascontiguousarray(tiles, dtype=np.uint8).reshape(-1)
        r = Result()
        self._check(lib().ua_puzzle_solve_bounded(self.h, n, t.ctypes.data, mem_bytes, C.byref(r)),
                    (UA_OK, UA_NO_PATH))
        return r.as_dict()

    def puzzle_moves(self) -> list:
        n = C.c_int64()
        self._check(lib().ua_puzzle_moves(self.h, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.int32)
        self._check(lib().ua_puzzle_moves(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value].tolist()

    # ---- batch
    @staticmethod
    def grid_query(grid: Grid, sx, sy, ex, ey) -> Query:
        q = Query()
        q.kind, q.grid, q.sx, q.sy, q.ex, q.ey = QUERY_GRID, grid.handle, sx, sy, ex, ey
        return q

    @staticmethod
    def puzzle_query(n, tiles) -> Query:
        q = Query()
        q.kind, q.n = QUERY_PUZZLE, n
        for i, v in enumerate(np.asarray(tiles).reshape(-1)):
            q.tiles[i] = int(v)
        return q

    def batch_solve(self, queries) -> list:
        arr = (Query * len(queries))(*queries)
        res = (Result * len(queries))()
        self._check(lib().ua_batch_solve(self.h, arr, len(queries), res))
        return [r.as_dict() for r in res]

    def batch_path(self, i: int):
        n = C.c_int64()
        self._check(lib().ua_batch_path(self.h, i, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.int32)
        self._check(lib().ua_batch_path(self.h, i, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    # ---- serving
    def serve_submit(self, q: Query) -> int:
        t = C.c_uint64()
        self._check(lib().ua_serve_submit(self.h, C.byref(q), C.byref(t)))
        return t.value

    def serve_step(self):
        a, p = C.c_int64(), C.c_int64()
        self._check(lib().ua_serve_step(self.h, C.byref(a), C.byref(p)))
        return a.value, p.value

    def serve_poll(self, cap=4096):
        tickets = (C.c_uint64 * cap)()
        res = (Result * cap)()
        n = C.c_int64()
        self._check(lib().ua_serve_poll(self.h, tickets, res, cap, C.byref(n)))
        return [(tickets[i], res[i].as_dict()) for i in range(n.value)]

    def serve_drain(self):
        self._check(lib().ua_serve_drain(self.h))

    def serve_path(self, ticket: int) -> np.ndarray:
        """Cell ids (grid) or blank moves (puzzle) of a served query that found its goal
        (contexts made with serve_paths=True); released once fetched."""
        n = C.c_int64()
        self._check(lib().ua_serve_path(self.h, C.c_uint64(ticket), None, 0, C.byref(n)))
        out = np.empty(max(1, n.value), np.int32)
        self._check(lib().ua_serve_path(self.h, C.c_uint64(ticket), out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def serve_stats(self):
        n, ms, rb, wd = C.c_int64(), C.c_double(), C.c_int32(), C.c_int64()
        self._check(lib().ua_serve_stats(self.h, C.byref(n), C.byref(ms), C.byref(rb), C.byref(wd)))
        return {"launches": n.value, "kernel_ms": ms.value, "budget_us": rb.value, "widened": wd.value}

    def launch_stats(self):
        """Every search launch so far: graph replays, direct launches, host wall and kernel ms."""
        g, d, w, k = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
        self._check(lib().ua_launch_stats(self.h, C.byref(g), C.byref(d), C.byref(w), C.byref(k)))
        return {"graph_launches": g.value, "direct_launches": d.value, "wall_ms": w.value, "kernel_ms": k.value,
                "graph_error": lib().ua_launch_graph_error(self.h).decode(errors="replace")}

    # ---- autotuning
    def autotune(self, queries) -> dict:
        """Measures candidate settings on a sample of queries and keeps the best per class."""
        arr = (Query * len(queries))(*queries)
        t = Tuning()
        self._check(lib().ua_autotune(self.h, arr, len(queries), C.byref(t)))
        return t.as_dict()

    def get_tuning(self) -> dict:
        t = Tuning()
        self._check(lib().ua_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, d: dict):
        t = Tuning()
        for k, v in d.items():
            setattr(t, k, v)
        self._check(lib().ua_set_tuning(self.h, C.byref(t)))

    # ---- diagnostics
    def atomic_rate(self, log2_words=26, blocks=4096, iters=64) -> float:
        """Random agent-scope atomicMin throughput of the device, in G atomics/s."""
        g = C.c_double()
        self._check(lib().ua_debug_atomic_rate(self.h, log2_words, blocks, iters, C.byref(g)))
        return g.value

    def access_latency(self, log2_words, kind, iters=2000) -> float:
        """ns per dependent access (kind 0 wg load, 1 wg atomicMin, 2 agent atomicMin, 3 agent load)."""
        ns = C.c_double()
        self._check(lib().ua_debug_access_latency(self.h, log2_words, kind, iters, C.byref(ns)))
        return ns.value

    def heap_selftest(self, ops, keys, cap_nodes):
        ops = np.ascontiguousarray(ops, dtype=np.int32)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        npops = int((ops < 0).sum())
        out = np.empty(max(npops * 64, 1), np.uint64)
        fs = C.c_int32()
        self._check(lib().ua_debug_heap_selftest(self.h, ops.ctypes.data, len(ops), keys.ctypes.data, keys.size,
                                                 cap_nodes, out.ctypes.data, out.size, C.byref(fs)))
        return out[: npops * 64].reshape(npops, 64), fs.value
