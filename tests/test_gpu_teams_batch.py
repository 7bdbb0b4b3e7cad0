"""GPU parity of batched solving on k_teams (one-block groups, one query per group,
pulled from a shared counter) against the oracle, and equality of one query's
result alone, in the plain whole-GPU launch and inside a mixed batch."""
import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path, goal_board, grid_bytes, random_grid, replay_moves

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def tctx():
    # far above what these queries need (the CPU model of the same code needs < 100 rounds)
    c = U.Context(0, max_rounds=50_000)
    yield c
    c.close()


def test_teams_batch_small(tctx, oracle):
    rng = np.random.default_rng(2024)
    ctx = tctx
    grids, queries, expect = [], [], []
    for i in range(8):
        H, W = int(rng.integers(3, 50)), int(rng.integers(3, 50))
        m = random_grid(rng, H, W, float(rng.uniform(0.05, 0.35)))
        sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
        m[sx, sy] = True
        g = grid_bytes(m)
        grid = ctx.grid_from_bytes(g, H, W)
        grids.append((grid, g, H, W, sx, sy, ex, ey))
        queries.append(ctx.grid_query(grid, sx, sy, ex, ey))
        expect.append(oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref"))
    _say(f"teams batch of {len(queries)} grid queries")
    res = ctx.batch_solve(queries)
    for i, (grid, g, H, W, sx, sy, ex, ey) in enumerate(grids):
        e, r = expect[i], res[i]
        assert r["status"] in (0, 1), (i, r)
        assert bool(r["found"]) == e["found"], (i, r, e)
        if e["found"]:
            assert (r["orth"], r["diag"]) == (e["orth"], e["diag"]), (i, r, e)
            check_grid_path(ctx.batch_path(i), g, H, W, sx, sy, ex, ey, r["orth"], r["diag"])
    for item in grids:
        item[0].free()


def test_teams_batch_mixed(tctx, oracle):
    ctx = tctx
    rng = np.random.default_rng(77)
    grids, queries, expect = [], [], []
    for i in range(60):
        H, W = int(rng.integers(2, 160)), int(rng.integers(2, 160))
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
    for s in range(16):
        b = oracle.puzzle_random(3, 7000 + s)
        boards.append(b)
        queries.append(ctx.puzzle_query(3, b))
        expect.append(oracle.puzzle_solve(3, b, pdb3, "astar")[0])
    _say(f"teams batch of {len(queries)} mixed queries")
    res = ctx.batch_solve(queries)
    for i, (grid, g, H, W, sx, sy, ex, ey) in enumerate(grids):
        e, r = expect[i], res[i]
        assert bool(r["found"]) == e["found"], (i, r, e)
        if e["found"]:
            assert (r["orth"], r["diag"]) == (e["orth"], e["diag"]), (i, r, e)
    for j, b in enumerate(boards):
        i = len(grids) + j
        assert res[i]["found"] and res[i]["steps"] == expect[i]
        mv = ctx.batch_path(i).tolist()
        assert len(mv) == expect[i] and replay_moves(3, b, mv) == goal_board(3)
    for item in grids:
        item[0].free()


def test_same_result_alone_plain_and_in_mix(tctx, oracle):
    ctx = tctx
    H = W = 400
    grid = ctx.grid_generate(H, W, 4242, 0.3, 0, 0, H - 1, W - 1)
    q = ctx.grid_query(grid, 0, 0, H - 1, W - 1)
    alone = ctx.batch_solve([q])[0]
    plain = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
    others, mix = [], [q]
    for k in range(24):
        s = 96 + 8 * k
        g2 = ctx.grid_generate(s, s, 500 + k, 0.25, 0, 0, s - 1, s - 1)
        others.append(g2)
        mix.append(ctx.grid_query(g2, 0, 0, s - 1, s - 1))
    mix.append(q)
    _say("alone / plain / in-mix")
    res = ctx.batch_solve(mix)
    g = oracle.grid_bernoulli(H, W, 4242, 0.3, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    pair = (ref["orth"], ref["diag"])
    for r in (alone, plain, res[0], res[-1]):
        assert r["found"] and (r["orth"], r["diag"]) == pair
    for g2 in others:
        g2.free()
    grid.free()
