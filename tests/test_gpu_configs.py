"""GPU parity at the sizes of the benchmark configurations:

- 16384 x 16384 grid, Bernoulli(0.35) from seed 2, opposite corners (plain whole-GPU
  solve) against the stored answer of the dense exact oracle;
- all 256 random solvable 15-puzzles of seeds 3000..3255, one plain launch each,
  against the stored CPU A* move counts;
- one batch of 4096 uneven grids plus 512 15-puzzles (seeds 41000..41511): every
  puzzle against the stored CPU A* answer, every grid against the oracle run here,
  every found grid cost above its octile bound, every path valid with exactly its
  reported cost.
"""
import concurrent.futures as cf
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import (c4_grid_specs, check_grid_path_bitmap, goal_board, replay_moves)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SQRT2 = 2 ** 0.5


def _say(msg):
    print(msg, flush=True)


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.mark.timeout(300)
def test_c2_grid_16384(ctx):
    row = [r for r in _golden("grid_configs.json") if r["H"] == 16384][0]
    H = W = row["H"]
    _say(f"C2 {H}x{W} seed {row['seed']} p {row['block_rate']}")
    grid = ctx.grid_generate(H, W, row["seed"], row["block_rate"], 0, 0, H - 1, W - 1)
    try:
        res = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
        _say(f"C2 gpu {res}")
        assert res["found"] == row["found"]
        assert (res["orth"], res["diag"]) == (row["orth"], row["diag"])
        path = ctx.grid_path()
        check_grid_path_bitmap(path, grid.download(), H, W, 0, 0, H - 1, W - 1, res["orth"], res["diag"])
    finally:
        grid.free()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("chunk", range(4))
def test_c3_plain_all(ctx, chunk):
    ctx.puzzle_build_pdb(4)
    rows = _golden("c3_puzzles.json")[chunk * 64:(chunk + 1) * 64]
    for r in rows:
        b = np.array(r["board"], np.uint8)
        res = ctx.puzzle_solve(4, b)
        assert res["found"] and res["steps"] == r["steps"], (r["seed"], res, r["steps"])
        mv = ctx.puzzle_moves()
        assert len(mv) == r["steps"] and replay_moves(4, b, mv) == goal_board(4), r["seed"]
    _say(f"C3 chunk {chunk}: {len(rows)} puzzles equal to the CPU A* move counts")


def _octile_pair(sx, sy, ex, ey):
    dx, dy = abs(sx - ex), abs(sy - ey)
    return abs(dx - dy), min(dx, dy)


@pytest.mark.timeout(900)
def test_c4_batch(oracle):
    gspecs = c4_grid_specs()
    pz_path = os.path.join(GOLDEN, "c4_puzzles.json")
    prows = _golden("c4_puzzles.json") if os.path.exists(pz_path) else \
        [{"seed": s, "board": oracle.puzzle_random(4, s).tolist(), "steps": None} for s in range(41000, 41512)]
    with U.Context(0, max_rounds=1_000_000) as ctx:
        ctx.puzzle_build_pdb(4)
        grids = [ctx.grid_generate(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"]) for s in gspecs]
        queries = [ctx.grid_query(g, *s["s"], *s["g"]) for g, s in zip(grids, gspecs)]
        queries += [ctx.puzzle_query(4, r["board"]) for r in prows]
        cells = sum(s["H"] * s["W"] for s in gspecs)
        _say(f"C4 batch: {len(gspecs)} grids ({cells / 1e9:.2f} G cells) + {len(prows)} puzzles")
        res = ctx.batch_solve(queries)
        gres, pres = res[:len(gspecs)], res[len(gspecs):]
        assert all(r["status"] in (0, 1) for r in res), [r for r in res if r["status"] not in (0, 1)][:3]
        _say(f"C4 batch solved: {sum(r['found'] for r in gres)} grids with a path")
        for r, p in zip(pres, prows):
            assert r["found"], p["seed"]
            if p["steps"] is not None:
                assert r["steps"] == p["steps"], (p["seed"], r["steps"], p["steps"])
        for r, s in zip(gres, gspecs):
            if r["found"]:
                lo, di = _octile_pair(*s["s"], *s["g"])
                assert r["orth"] + SQRT2 * r["diag"] >= lo + SQRT2 * di - 1e-9
        # every grid against the oracle's dense exact A*, on 16 host threads
        pick = list(range(len(gspecs)))

        def solve(i):
            s = gspecs[i]
            g = oracle.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
            return i, oracle.grid_astar(g, s["H"], s["W"], *s["s"], *s["g"], "dense", want_path=False)

        with cf.ThreadPoolExecutor(16) as ex:
            for i, e in ex.map(solve, pick):
                r = gres[i]
                assert bool(r["found"]) == e["found"], (i, gspecs[i], r, e)
                if e["found"]:
                    assert (r["orth"], r["diag"]) == (e["orth"], e["diag"]), (i, gspecs[i], r, e)
        _say(f"C4: {len(pick)} grids equal to the oracle")
        # Paths: the batch runs in memory-bounded sub-batches; a sub-batch's paths are walked
        # before its state is freed, so every found query has one. Every grid's path is
        # checked: legal moves out of open cells, with exactly the reported step counts
        checked = 0
        for i, (s, r) in enumerate(zip(gspecs, gres)):
            if not r["found"]:
                continue
            path = ctx.batch_path(i)
            check_grid_path_bitmap(path, grids[i].download(), s["H"], s["W"], *s["s"], *s["g"], r["orth"], r["diag"])
            checked += 1
        for i, p in enumerate(prows):
            mv = ctx.batch_path(len(gspecs) + i)
            assert len(mv) == pres[i]["steps"]
            assert replay_moves(4, np.array(p["board"], np.uint8), mv) == goal_board(4)
            checked += 1
        _say(f"C4: {checked} paths / move lists valid")
        assert checked == sum(r["found"] for r in gres) + len(prows)
        for g in grids:
            g.free()
