"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py pdb      # PDB digests (sha256 + depth histogram)
    python tests/golden/make_golden.py c3       # 256 random solvable 15-puzzles, seeds 3000..3255
    python tests/golden/make_golden.py c4       # 512 random solvable 15-puzzles, seeds 41000..41511
    python tests/golden/make_golden.py grids    # seeded grid configs (1000^2 seed 1; 16384^2 seed 2)

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


def cmd_pdb():
    out = {}
    for n in (3, 4):
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
    with open(os.path.join(HERE, "pdb_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


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


if __name__ == "__main__":
    {"pdb": cmd_pdb, "c3": cmd_c3, "c4": cmd_c4, "grids": cmd_grids}[sys.argv[1]]()
