"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py pdb      # PDB digests (sha256 + depth histogram)
    python tests/golden/make_golden.py c3       # 256 random solvable 15-puzzles, seeds 3000..3255
    python tests/golden/make_golden.py c4       # 512 random solvable 15-puzzles, seeds 41000..41511
    python tests/golden/make_golden.py grids    # seeded grid configs (1000^2 seed 1; 16384^2 seed 2)
    python tests/golden/make_golden.py pdb5     # the four 6-tile 5x5 PDB digests (added to pdb_digests.json)
    python tests/golden/make_golden.py c5       # 5x5 boards from seeded random walks off the goal
    python tests/golden/make_golden.py c5stream # queries of the serving stream (seed 5) with dense-oracle answers

The oracle itself is pinned by tests/test_oracle.py (closed forms, exhaustive BFS,
textbook Dijkstra / IDA*); these files record its answers at sizes the GPU tests
cannot recompute quickly.
"""
from __future__ import annotations

import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

CACHE = os.path.join(ROOT, "_cache")


def cmd_pdb(sizes=(3, 4)):
    path = os.path.join(HERE, "pdb_digests.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for n in sizes:
        pdb = O.pdb_full(n, 1, cache_dir=CACHE)
        parts = O.partitions(n)
        off = 0
        tabs = []
        for t in parts:
            size = O.pdb_size(n, len(t))
            tab = pdb[off: off + size]
            off += size
            tabs.append({"tiles": t, "size": int(size), "sha256": hashlib.sha256(tab.tobytes()).hexdigest(),
                         "histogram": np.bincount(tab, minlength=256).tolist()})
        out[str(n)] = tabs
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


def walk_board(n, seed, length):
    """The board `length` blank moves away from the goal along a seeded walk that never
    undoes its previous move (moves drawn from the oracle's mix64 stream)."""
    b = list(range(1, n * n)) + [0]
    z, last, s = n * n - 1, -1, seed
    for _ in range(length):
        while True:
            s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
            d = O.lib().orc_mix64(s) % 4
            x, y = z // n + (1, -1, 0, 0)[d], z % n + (0, 0, 1, -1)[d]
            if 0 <= x < n and 0 <= y < n and (last < 0 or d != (last ^ 1)):
                break
        nz = x * n + y
        b[z], b[nz] = b[nz], b[z]
        z, last = nz, d
    return b


def cmd_c5():
    """5x5 boards from walks of 30..120 moves; optimal steps by IDA* with Manhattan
    distance (textbook) and by A* with the reference's 6-6-6-6 PDB, which must agree."""
    pdb5 = O.pdb_full(5, 1, cache_dir=CACHE)
    rows = []
    for i, length in enumerate([30, 40, 50, 60, 70, 80, 90, 100, 110, 120] * 2):
        seed = 5000 + i
        b = walk_board(5, seed, length)
        t = time.time()
        steps, moves, exp = O.puzzle_solve(5, np.array(b, np.uint8), None, "ida_manhattan", node_limit=400_000_000)
        t_ida = time.time() - t
        if steps < 0:
            print(seed, length, "IDA* node limit", flush=True)
            continue
        t = time.time()
        s2, _, exp2 = O.puzzle_solve(5, np.array(b, np.uint8), pdb5, "astar", node_limit=20_000_000)
        assert s2 == steps, (seed, s2, steps)
        rows.append({"seed": seed, "walk": length, "board": b, "steps": steps, "moves": moves,
                     "expanded": exp2, "ida_nodes": exp, "cpu_s": round(time.time() - t, 3)})
        print(seed, length, steps, exp, round(t_ida, 1), exp2, flush=True)
    with open(os.path.join(HERE, "c5_puzzles.json"), "w") as f:
        json.dump(rows, f)


_pdb4 = None


def _solve_seed(seed):
    global _pdb4
    if _pdb4 is None:
        _pdb4 = O.pdb_full(4, 1, cache_dir=CACHE)
    b = O.puzzle_random(4, seed)
    t = time.time()
    steps, moves, exp = O.puzzle_solve(4, b, _pdb4, "astar")
    return {"seed": seed, "board": b.tolist(), "steps": steps, "moves": moves, "expanded": exp,
            "cpu_s": round(time.time() - t, 3)}


def cmd_c3(workers=6, seeds=None, out_name="c3_puzzles.json"):
    O.pdb_full(4, 1, cache_dir=CACHE)
    seeds = seeds or list(range(3000, 3256))
    with mp.Pool(workers) as pool:
        rows = []
        for r in pool.imap_unordered(_solve_seed, seeds):
            rows.append(r)
            print(r["seed"], r["steps"], r["expanded"], r["cpu_s"], flush=True)
    rows.sort(key=lambda r: r["seed"])
    with open(os.path.join(HERE, out_name), "w") as f:
        json.dump(rows, f)


def cmd_c4(workers=5):
    """The 512 puzzles of the batched configuration: boards from seeds 41000..41511."""
    cmd_c3(workers, list(range(41000, 41512)), "c4_puzzles.json")


def cmd_grids():
    rows = []
    for (H, seed, p, variant) in [(1000, 1, 0.30, "ref"), (16384, 2, 0.35, "dense")]:
        g = O.grid_bernoulli(H, H, seed, p, 0, 0, H - 1, H - 1)
        t = time.time()
        r = O.grid_astar(g, H, H, 0, 0, H - 1, H - 1, variant, want_path=False)
        rows.append({"H": H, "W": H, "seed": seed, "block_rate": p, "start": [0, 0], "goal": [H - 1, H - 1],
                     "found": r["found"], "orth": r["orth"], "diag": r["diag"], "expanded": r["expanded"],
                     "variant": variant, "cpu_s": round(time.time() - t, 2)})
        print(rows[-1], flush=True)
        del g
    with open(os.path.join(HERE, "grid_configs.json"), "w") as f:
        json.dump(rows, f, indent=1)


def _solve_stream_grid(s):
    g = O.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
    t = time.time()
    r = O.grid_astar(g, s["H"], s["W"], *s["s"], *s["g"], "dense", want_path=False)
    return {"found": bool(r["found"]), "orth": int(r["orth"]), "diag": int(r["diag"]),
            "expanded": int(r["expanded"]), "cpu_s": round(time.time() - t, 2)}


def cmd_c5stream(workers=6, first=40, deepest=3):
    """Queries of the serving stream of the bench (bench.mixed_stream, seed 5, 204 queries
    in stream order: 70 % grids with side log-uniform in [512, 8192], 30 % 15-puzzles
    from the c3 boards): the first `first` of them plus the `deepest` grids of the
    whole stream (largest start-goal Chebyshev distance). Grid answers come from the
    oracle's dense exact A*; puzzle answers are the c3 fixture's."""
    import bench
    specs = bench.mixed_stream(204)
    grids = [i for i, s in enumerate(specs) if s["kind"] == "grid"]
    cheb = lambda s: max(abs(s["s"][0] - s["g"][0]), abs(s["s"][1] - s["g"][1]))
    deep = sorted(grids, key=lambda i: -cheb(specs[i]))[:deepest]
    pick = sorted(set(range(first)) | set(deep))
    todo = [specs[i] for i in pick if specs[i]["kind"] == "grid"]
    with mp.Pool(workers) as pool:
        answers = pool.map(_solve_stream_grid, todo, chunksize=1)
    it = iter(answers)
    rows = []
    for i in pick:
        s = specs[i]
        if s["kind"] == "grid":
            a = next(it)
            rows.append({"index": i, "kind": "grid", "H": s["H"], "W": s["W"], "seed": s["seed"], "p": s["p"],
                         "start": list(s["s"]), "goal": list(s["g"]), "deep": i in deep, **a})
        else:
            rows.append({"index": i, "kind": "puzzle", "board": s["board"], "steps": s["steps"],
                         "expanded": s["expanded"]})
        print(rows[-1] if s["kind"] == "puzzle" else {k: v for k, v in rows[-1].items() if k != "seed"}, flush=True)
    with open(os.path.join(HERE, "c5_stream.json"), "w") as f:
        json.dump(rows, f, indent=0)


if __name__ == "__main__":
    {"pdb": cmd_pdb, "pdb5": lambda: cmd_pdb((5,)), "c3": cmd_c3, "c4": cmd_c4, "grids": cmd_grids,
     "c5": cmd_c5, "c5stream": cmd_c5stream}[sys.argv[1]]()
