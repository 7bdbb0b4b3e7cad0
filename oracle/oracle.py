"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests/, the smoke check in __graft_entry__.py and the
cpu_baseline leg of bench.py import this module as the checker. The product
(uastar_amd/, include/, the CLI) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


class GridOut(C.Structure):
    _fields_ = [
        ("found", C.c_int32),
        ("orth", C.c_int64),
        ("diag", C.c_int64),
        ("cost_f32", C.c_float),
        ("expanded", C.c_int64),
        ("generated", C.c_int64),
        ("inserted", C.c_int64),
        ("path_len", C.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u8p = C.POINTER(C.c_uint8)
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_cell_hash.restype = C.c_uint64
        L.orc_cell_hash.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_grid_bernoulli.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                         C.c_int, C.c_int, C.c_int, C.c_int, u8p]
        L.orc_grid_zigzag.argtypes = [C.c_int, C.c_int, C.c_int, u8p] + [C.POINTER(C.c_int)] * 4
        L.orc_puzzle_random.argtypes = [C.c_int, C.c_uint64, u8p]
        L.orc_puzzle_solvable.argtypes = [C.c_int, u8p]
        L.orc_grid_astar.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.POINTER(GridOut), C.POINTER(C.c_int32), C.c_int64]
        L.orc_pdb_size.restype = C.c_uint64
        L.orc_pdb_size.argtypes = [C.c_int, C.c_int]
        L.orc_pdb_encode.restype = C.c_uint32
        L.orc_pdb_encode.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, u8p]
        L.orc_pdb_decode.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_uint32, u8p]
        L.orc_pdb_build.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, u8p]
        L.orc_partitions.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_puzzle_astar.argtypes = [C.c_int, u8p, u8p, C.c_int, C.c_int64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int64,
                                       C.POINTER(C.c_int64)]
        L.orc_puzzle_h.argtypes = [C.c_int, u8p, u8p]
        _lib = L
    return _lib


def _u8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _ints(xs):
    arr = (C.c_int * len(xs))(*xs)
    return arr


# ---------------------------------------------------------------- generators

def cell_hash(seed: int, cid: int) -> int:
    return lib().orc_cell_hash(seed, cid)


def block_threshold(p: float) -> int:
    return int(np.floor(p * 4294967296.0))


def grid_bernoulli(H, W, seed, p, sx, sy, ex, ey) -> np.ndarray:
    g = np.empty(H * W, np.uint8)
    lib().orc_grid_bernoulli(H, W, seed, block_threshold(p), sx, sy, ex, ey, _u8(g))
    return g


def grid_zigzag(H, W, gap):
    g = np.empty(H * W, np.uint8)
    v = [C.c_int() for _ in range(4)]
    lib().orc_grid_zigzag(H, W, gap, _u8(g), *[C.byref(x) for x in v])
    return g, tuple(x.value for x in v)


def puzzle_random(n, seed) -> np.ndarray:
    t = np.empty(n * n, np.uint8)
    lib().orc_puzzle_random(n, seed, _u8(t))
    return t


def puzzle_solvable(n, tiles) -> bool:
    t = np.ascontiguousarray(tiles, dtype=np.uint8)
    return bool(lib().orc_puzzle_solvable(n, _u8(t)))


# ---------------------------------------------------------------- grid search

VARIANT = {"ref": 0, "dense": 1, "f32": 2, "dijkstra": 3}


def grid_astar(graph: np.ndarray, H, W, sx, sy, ex, ey, variant="ref", want_path=True):
    """Returns dict(found, orth, diag, cost_f32, expanded, generated, inserted, path)."""
    g = np.ascontiguousarray(graph, dtype=np.uint8).reshape(-1)
    assert g.size == H * W
    out = GridOut()
    cap = H * W if want_path else 0
    path = np.empty(max(cap, 1), np.int32)
    rc = lib().orc_grid_astar(_u8(g), H, W, sx, sy, ex, ey, VARIANT[variant], C.byref(out),
                              path.ctypes.data_as(C.POINTER(C.c_int32)) if want_path else None, cap)
    if rc != 0:
        raise ValueError(f"orc_grid_astar rc={rc}")
    res = {f: getattr(out, f) for f, _ in GridOut._fields_}
    res["found"] = bool(res["found"])
    res["path"] = path[: out.path_len].copy() if (want_path and out.found) else None
    return res


# ---------------------------------------------------------------- PDB / puzzle

def partitions(n):
    counts = (C.c_int * 4)()
    tiles = (C.c_int * 32)()
    c = lib().orc_partitions(n, counts, tiles)
    return [[tiles[i * 8 + j] for j in range(counts[i])] for i in range(c)]


def pdb_size(n, k):
    return lib().orc_pdb_size(n, k)


def pdb_encode(n, tracked, conf) -> int:
    c = np.ascontiguousarray(conf, dtype=np.uint8)
    return lib().orc_pdb_encode(n, _ints(tracked), len(tracked), _u8(c))


def pdb_decode(n, tracked, code) -> np.ndarray:
    out = np.zeros(n * n, np.uint8)
    lib().orc_pdb_decode(n, _ints(tracked), len(tracked), code, _u8(out))
    return out


def pdb_build(n, tracked, method=1) -> np.ndarray:
    out = np.empty(pdb_size(n, len(tracked)), np.uint8)
    rc = lib().orc_pdb_build(n, _ints(tracked), len(tracked), method, _u8(out))
    if rc:
        raise ValueError("bad pattern")
    return out


def pdb_full(n, method=1, cache_dir=None) -> np.ndarray:
    """All partition tables of the reference, concatenated (GPU-solver.cuh:459-467).
    With cache_dir, each table is kept as database_<n>_<tiles>.bin (database.cpp:88-91)."""
    parts = []
    for t in partitions(n):
        path = None
        if cache_dir:
            os.makedirs(cache_dir, exist_ok=True)
            path = os.path.join(cache_dir, "database_%d_%s.bin" % (n, "_".join(map(str, t))))
            if os.path.exists(path) and os.path.getsize(path) == pdb_size(n, len(t)):
                parts.append(np.fromfile(path, np.uint8))
                continue
        tab = pdb_build(n, t, method)
        if path:
            tab.tofile(path + ".tmp")
            os.replace(path + ".tmp", path)
        parts.append(tab)
    return np.concatenate(parts)


PUZZLE_VARIANT = {"astar": 0, "ida_pdb": 1, "ida_manhattan": 2}


def puzzle_solve(n, tiles, pdb=None, variant="astar", node_limit=0):
    """Returns (steps or -1 unsolvable / -3 limit, moves list, expanded)."""
    t = np.ascontiguousarray(tiles, dtype=np.uint8)
    exp = C.c_int64()
    ml = C.c_int64()
    moves = np.empty(4096, np.int32)
    p = _u8(pdb) if pdb is not None else None
    rc = lib().orc_puzzle_astar(n, _u8(t), p, PUZZLE_VARIANT[variant], node_limit, C.byref(exp),
                                moves.ctypes.data_as(C.POINTER(C.c_int32)), moves.size, C.byref(ml))
    if rc == -2:
        raise ValueError("bad puzzle arguments")
    mv = moves[: ml.value].tolist() if rc >= 0 else None
    return rc, mv, exp.value


def puzzle_h(n, tiles, pdb) -> int:
    t = np.ascontiguousarray(tiles, dtype=np.uint8)
    return lib().orc_puzzle_h(n, _u8(t), _u8(pdb))
