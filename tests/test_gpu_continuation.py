"""GPU tests of continuation inside the round (grid searches).

A one-block grid team whose heaps keep keys after the pop expands a lane's equal-f child
in the same round. Without it, a round advances a path by at most one step, so a solve
takes more rounds than its path has steps. These tests check three things:
- with continuation, large depth-bound grids finish in fewer rounds than their path has
  steps;
- the answers still equal the oracle's exact pair;
- the returned path is valid.
A wide plain team, whose heaps run dry each round, must not multiply its expansions."""
import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path_bitmap

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("seed,p", [(11, 0.20), (12, 0.30)])
def test_depth_bound_grid_takes_fewer_rounds_than_steps(ctx, oracle, seed, p):
    H = W = 2000
    _say(f"continuation grid {H}x{W} seed {seed} p {p}")
    e = oracle.grid_astar(oracle.grid_bernoulli(H, W, seed, p, 0, 0, H - 1, W - 1), H, W, 0, 0, H - 1, W - 1,
                          "ref", want_path=False)
    grid = ctx.grid_generate(H, W, seed, p, 0, 0, H - 1, W - 1)
    try:
        r = ctx.batch_solve([ctx.grid_query(grid, 0, 0, H - 1, W - 1)])[0]
        _say(f"batch: {r}")
        assert r["status"] == 0 and bool(r["found"]) == e["found"]
        if e["found"]:
            assert (r["orth"], r["diag"]) == (e["orth"], e["diag"])
            assert r["rounds"] < r["orth"] + r["diag"], "no round advanced a chain by two steps"
        p1 = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
        assert (p1["found"], p1["orth"], p1["diag"]) == (r["found"], r["orth"], r["diag"])
        if e["found"]:
            check_grid_path_bitmap(ctx.grid_path(), grid.download(), H, W, 0, 0, H - 1, W - 1, p1["orth"], p1["diag"])
    finally:
        grid.free()


@pytest.mark.timeout(300)
def test_wide_plain_team_does_not_multiply_expansions():
    """1024 heaps on a 1000^2 grid: the whole open list fits in one round's pops, so no
    wave continues. The expansions stay near the cell count instead of multiplying."""
    with U.Context(0, plain_heaps=1024) as c:
        H = W = 1000
        g = c.grid_generate(H, W, 1, 0.30, 0, 0, H - 1, W - 1)
        try:
            r = c.grid_solve(g, 0, 0, H - 1, W - 1)
            _say(f"plain 1024 heaps: {r}")
            assert r["status"] == 0 and r["found"]
            assert r["expanded"] < 2 * H * W, r["expanded"]
        finally:
            g.free()
