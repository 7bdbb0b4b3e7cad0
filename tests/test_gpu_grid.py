"""GPU parity of the grid GA* path against the CPU oracle (exact (orth, diag) costs).

Bar: found flags and cost pairs bit-identical to the oracle; every returned path valid.
"""
import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path, grid_bytes, random_grid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_heap_selftest(ctx):
    rng = np.random.default_rng(7)
    for trial in range(6):
        ops, keys, ref, pops = [], [], [], []
        for _ in range(400):
            if rng.random() < 0.6:
                c = int(rng.integers(0, 65))
                ks = rng.integers(0, [10, 1000, 1 << 62][trial % 3], size=c, dtype=np.uint64)
                ops.append(c)
                keys.extend(ks.tolist())
                ref.extend(ks.tolist())
            else:
                ops.append(-1)
                ref.sort()
                pops.append(ref[:64])
                del ref[:64]
        _say(f"heap trial {trial}: {len(ops)} ops")
        got, size = ctx.heap_selftest(ops, np.array(keys, np.uint64), cap_nodes=4096)
        assert size == len(ref)
        for i, exp in enumerate(pops):
            row = [int(v) for v in got[i] if int(v) != (1 << 64) - 1]
            assert row == [int(v) for v in exp], f"trial {trial} pop {i}"


def test_heap_overflow_reported(ctx):
    ops = [64] * 8
    keys = np.arange(512, dtype=np.uint64)[::-1].copy()
    _, size = ctx.heap_selftest(ops, keys, cap_nodes=3)
    assert size == -1


def _solve_both(ctx, oracle, open_mask, sx, sy, ex, ey):
    H, W = open_mask.shape
    g = grid_bytes(open_mask)
    ref = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
    grid = ctx.grid_from_bytes(g, H, W)
    try:
        res = ctx.grid_solve(grid, sx, sy, ex, ey)
        path = ctx.grid_path() if res["found"] else None
    finally:
        grid.free()
    return g, ref, res, path


def _assert_parity(g, H, W, sx, sy, ex, ey, ref, res, path):
    assert bool(res["found"]) == ref["found"]
    if ref["found"]:
        assert (res["orth"], res["diag"]) == (ref["orth"], ref["diag"])
        check_grid_path(path, g, H, W, sx, sy, ex, ey, res["orth"], res["diag"])
    else:
        assert res["status"] == U.UA_NO_PATH


def test_readme_example(ctx, oracle):
    # main.cpp:19-26: expected (4 orth, 3 diag) = 4 + 3*sqrt(2)
    m = np.array([[1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 1, 1, 0, 1], [1, 1, 1, 0, 1]])
    g, ref, res, path = _solve_both(ctx, oracle, m, 0, 0, 4, 4)
    assert (res["orth"], res["diag"]) == (4, 3)
    _assert_parity(g, 5, 5, 0, 0, 4, 4, ref, res, path)


def test_edge_cases(ctx, oracle):
    cases = []
    m = np.ones((7, 9), bool)
    cases.append((m, 3, 3, 3, 3))                      # start == goal
    m2 = m.copy(); m2[0, 0] = False
    cases.append((m2, 0, 0, 6, 8))                      # blocked start: no moves
    m3 = m.copy(); m3[1, :3] = False; m3[:2, 2] = False
    cases.append((m3, 0, 0, 6, 8))                      # enclosed start: no path
    m4 = m.copy(); m4[6, 8] = False
    cases.append((m4, 0, 0, 6, 8))                      # blocked goal is still reachable
    cases.append((np.ones((1, 40), bool), 0, 0, 0, 39))  # a single row
    cases.append((np.ones((40, 1), bool), 39, 0, 0, 0))  # a single column
    m5 = np.zeros((6, 6), bool); m5[0, 0] = True
    cases.append((m5, 0, 0, 1, 1))                      # goal next to an open start, all else blocked
    for i, (mm, sx, sy, ex, ey) in enumerate(cases):
        _say(f"edge case {i}: {mm.shape} ({sx},{sy})->({ex},{ey})")
        H, W = mm.shape
        g, ref, res, path = _solve_both(ctx, oracle, mm, sx, sy, ex, ey)
        _assert_parity(g, H, W, sx, sy, ex, ey, ref, res, path)


def test_partial_out_move_masks_are_rejected(ctx):
    """A cell byte other than 0 and 0xFF gates single moves (pathway/CPU-solver.cpp:64-66),
    which the 1-bit grid cannot hold: ua_grid_from_bytes refuses the grid (UA_ERR_INVALID)
    instead of treating the cell as fully open; 0 / 0xFF bytes still load."""
    H, W = 9, 70
    g = np.full(H * W, 0xFF, np.uint8)
    g[3 * W + 66] = 0x0F
    _say("partial mask 0x0F at (3, 66)")
    with pytest.raises(U.UAError) as e:
        ctx.grid_from_bytes(g, H, W)
    assert e.value.status == -1 and "partial out-move mask" in str(e.value)
    g[3 * W + 66] = 0x80
    with pytest.raises(U.UAError):
        ctx.grid_from_bytes(g, H, W)
    g[3 * W + 66] = 0
    grid = ctx.grid_from_bytes(g, H, W)
    r = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
    grid.free()
    assert r["found"] and (r["orth"], r["diag"]) == (61, 8)


@pytest.fixture(scope="module")
def team_ctxs():
    """Contexts whose plain solves run the other open lists and team shapes: one block on
    the bucket ring (16 heaps), four blocks on per-wave heaps (64), one block on per-wave
    heaps (UA_FLAG_GRID_HEAPS)."""
    cs = {"ring-1-block": U.Context(0, plain_heaps=16), "heaps-4-blocks": U.Context(0, plain_heaps=64),
          "heaps-1-block": U.Context(0, plain_heaps=16, grid_heaps=True)}
    yield cs
    for c in cs.values():
        c.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_small_every_team_shape(team_ctxs, oracle, seed):
    rng = np.random.default_rng(900 + seed)
    H, W = int(rng.integers(2, 150)), int(rng.integers(2, 150))
    m = random_grid(rng, H, W, float(rng.uniform(0.0, 0.45)))
    sx, sy, ex, ey = int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W))
    m[sx, sy] = True
    for name, c in team_ctxs.items():
        _say(f"random grid seed {seed}: {H}x{W} on {name}")
        g, ref, res, path = _solve_both(c, oracle, m, sx, sy, ex, ey)
        _assert_parity(g, H, W, sx, sy, ex, ey, ref, res, path)


@pytest.mark.parametrize("seed", range(12))
def test_random_small(ctx, oracle, seed):
    rng = np.random.default_rng(100 + seed)
    H, W = int(rng.integers(2, 80)), int(rng.integers(2, 80))
    p = float(rng.uniform(0.0, 0.45))
    m = random_grid(rng, H, W, p)
    sx, sy, ex, ey = int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W))
    m[sx, sy] = True
    _say(f"random grid seed {seed}: {H}x{W} p={p:.2f}")
    g, ref, res, path = _solve_both(ctx, oracle, m, sx, sy, ex, ey)
    _assert_parity(g, H, W, sx, sy, ex, ey, ref, res, path)


def test_generator_matches_oracle_spec(ctx, oracle):
    H, W = 131, 200
    grid = ctx.grid_generate(H, W, 12345, 0.3, 0, 0, H - 1, W - 1)
    try:
        dev = grid.to_bytes()
    finally:
        grid.free()
    host = oracle.grid_bernoulli(H, W, 12345, 0.3, 0, 0, H - 1, W - 1)
    assert np.array_equal(dev, host)


def test_c1_grid_1000(ctx, oracle):
    """1000x1000, Bernoulli(0.30) from seed 1, corners open."""
    H = W = 1000
    _say("C1 1000x1000 seed 1")
    g = oracle.grid_bernoulli(H, W, 1, 0.30, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    grid = ctx.grid_generate(H, W, 1, 0.30, 0, 0, H - 1, W - 1)
    try:
        res = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
        path = ctx.grid_path()
    finally:
        grid.free()
    _say(f"C1 gpu {res}")
    _assert_parity(g, H, W, 0, 0, H - 1, W - 1, ref, res, path)
