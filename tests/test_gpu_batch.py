"""GPU parity of the persistent batch launch (one query per workgroup, uneven sizes)
and the "same result alone and in the mix" property."""
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path, goal_board, grid_bytes, random_grid, replay_moves

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def bctx():
    # a round limit far above what these queries need (< 200 rounds each on the CPU model)
    c = U.Context(0, max_rounds=200_000)
    yield c
    c.close()


@pytest.mark.skip(reason="hung twice on the GPU with the former workgroup-team batch kernel (k_batch, removed); "
                         "cause not shown: the SIMT emulator runs the same inputs correctly. Kept off the GPU until "
                         "the cause is shown; the one-block-group batch path is covered by test_gpu_teams_batch.py")
def test_mixed_batch_matches_oracle(bctx, oracle):
    ctx = bctx
    rng = np.random.default_rng(42)
    grids, queries, expect = [], [], []
    for i in range(40):
        H, W = int(rng.integers(2, 120)), int(rng.integers(2, 120))
        m = random_grid(rng, H, W, float(rng.uniform(0.0, 0.4)))
        sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
        m[sx, sy] = True
        g = grid_bytes(m)
        grid = ctx.grid_from_bytes(g, H, W)
        grids.append((grid, g, H, W, sx, sy, ex, ey))
        queries.append(ctx.grid_query(grid, sx, sy, ex, ey))
        expect.append(oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref"))
    pdb3 = oracle.pdb_full(3)
    boards = []
    for s in range(12):
        b = oracle.puzzle_random(3, 900 + s)
        boards.append(b)
        queries.append(ctx.puzzle_query(3, b))
        expect.append(oracle.puzzle_solve(3, b, pdb3, "astar")[0])
    _say(f"batch of {len(queries)} queries")
    res = ctx.batch_solve(queries)
    for i, (grid, g, H, W, sx, sy, ex, ey) in enumerate(grids):
        e, r = expect[i], res[i]
        assert bool(r["found"]) == e["found"], (i, r, e)
        if e["found"]:
            assert (r["orth"], r["diag"]) == (e["orth"], e["diag"]), (i, r, e)
            check_grid_path(ctx.batch_path(i), g, H, W, sx, sy, ex, ey, r["orth"], r["diag"])
    for j, b in enumerate(boards):
        i = len(grids) + j
        assert res[i]["found"] and res[i]["steps"] == expect[i]
        mv = ctx.batch_path(i).tolist()
        assert len(mv) == expect[i] and replay_moves(3, b, mv) == goal_board(3)
    for item in grids:
        item[0].free()


@pytest.mark.skip(reason="hung twice on the GPU with the former workgroup-team batch kernel (k_batch, removed); "
                         "cause not shown: the SIMT emulator runs the same inputs correctly. Kept off the GPU until "
                         "the cause is shown; the one-block-group batch path is covered by test_gpu_teams_batch.py")
def test_alone_equals_in_mix(bctx, oracle):
    ctx = bctx
    """The same query gives the same cost alone, in a small batch and in a large mixed batch."""
    H = W = 300
    grid = ctx.grid_generate(H, W, 77, 0.3, 0, 0, H - 1, W - 1)
    q = ctx.grid_query(grid, 0, 0, H - 1, W - 1)
    alone = ctx.batch_solve([q])[0]
    plain = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
    rows = json.load(open(os.path.join(GOLDEN, "c3_puzzles.json")))
    easy = [r for r in rows if r["expanded"] < 400_000][:6]
    mix = [q]
    others = []
    for k in range(30):
        g2 = ctx.grid_generate(128 + 8 * k, 128 + 8 * k, 1000 + k, 0.25, 0, 0, 127 + 8 * k, 127 + 8 * k)
        others.append(g2)
        mix.append(ctx.grid_query(g2, 0, 0, 127 + 8 * k, 127 + 8 * k))
    for r in easy:
        mix.append(ctx.puzzle_query(4, r["board"]))
    mix.append(q)
    res = ctx.batch_solve(mix)
    g = oracle.grid_bernoulli(H, W, 77, 0.3, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    pair = (ref["orth"], ref["diag"])
    assert (alone["orth"], alone["diag"]) == pair
    assert (plain["orth"], plain["diag"]) == pair
    assert (res[0]["orth"], res[0]["diag"]) == pair and (res[-1]["orth"], res[-1]["diag"]) == pair
    for k, r in enumerate(easy):
        assert res[31 + k]["steps"] == r["steps"]
    for g2 in others:
        g2.free()
    grid.free()
