"""GPU tests of the serving loop (continuous batching): results equal the oracle's,
queries submitted while others are in flight are admitted and overlap, and a query
gives the same answer alone and in the mix."""
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import grid_bytes, random_grid

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sctx():
    c = U.Context(0, max_rounds=200_000, mem_budget=8 << 30)
    yield c
    c.close()


def _drain(ctx, want):
    got = {}
    for _ in range(100000):
        a, p = ctx.serve_step()
        for t, r in ctx.serve_poll():
            got[t] = r
        if len(got) >= want and a == 0 and p == 0:
            break
    return got


def test_serve_matches_oracle_with_late_arrivals(sctx, oracle):
    ctx = sctx
    rng = np.random.default_rng(5)
    grids, tickets, expect = [], {}, {}
    for i in range(24):
        H, W = int(rng.integers(16, 200)), int(rng.integers(16, 200))
        m = random_grid(rng, H, W, float(rng.uniform(0.1, 0.4)))
        sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
        m[sx, sy] = True
        g = grid_bytes(m)
        grid = ctx.grid_from_bytes(g, H, W)
        grids.append(grid)
        e = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
        q = ctx.grid_query(grid, sx, sy, ex, ey)
        expect_i = ("grid", e)
        if i < 12:
            tickets[ctx.serve_submit(q)] = expect_i
        else:
            expect[i] = (q, expect_i)
    rows = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))) if r["expanded"] < 300_000][:6]
    late = list(expect.values())
    _say("serving: 12 grids submitted, then steps with late arrivals")
    done = {}
    for step in range(6):  # first queries are in flight while these arrive
        ctx.serve_step()
        for t, r in ctx.serve_poll():
            done[t] = r
        for q, e in late[step * 2: step * 2 + 2]:
            tickets[ctx.serve_submit(q)] = e
    for r in rows:
        tickets[ctx.serve_submit(ctx.puzzle_query(4, r["board"]))] = ("puzzle", r["steps"])
    done.update(_drain(ctx, len(tickets) - len(done)))
    assert set(done) == set(tickets)
    for t, (kind, e) in tickets.items():
        r = done[t]
        if kind == "grid":
            assert bool(r["found"]) == e["found"], (t, r, e)
            if e["found"]:
                assert (r["orth"], r["diag"]) == (e["orth"], e["diag"]), (t, r, e)
        else:
            assert r["found"] and r["steps"] == e, (t, r, e)
        assert r["latency_ms"] > 0
    for g in grids:
        g.free()


def test_serve_same_alone_and_in_mix(sctx, oracle):
    ctx = sctx
    H = W = 500
    grid = ctx.grid_generate(H, W, 99, 0.3, 0, 0, H - 1, W - 1)
    q = ctx.grid_query(grid, 0, 0, H - 1, W - 1)
    t_alone = ctx.serve_submit(q)
    alone = _drain(ctx, 1)[t_alone]
    others = []
    ts = []
    for k in range(40):
        s = 100 + 10 * k
        g2 = ctx.grid_generate(s, s, 800 + k, 0.3, 0, 0, s - 1, s - 1)
        others.append(g2)
        ts.append(ctx.serve_submit(ctx.grid_query(g2, 0, 0, s - 1, s - 1)))
        if k == 20:
            t_mix = ctx.serve_submit(q)
    got = _drain(ctx, len(ts) + 1)
    g = oracle.grid_bernoulli(H, W, 99, 0.3, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    pair = (ref["orth"], ref["diag"])
    assert (alone["orth"], alone["diag"]) == pair
    assert (got[t_mix]["orth"], got[t_mix]["diag"]) == pair
    for g2 in others:
        g2.free()
    grid.free()


def test_serve_widens_long_puzzle(sctx):
    """Two heavy 15-puzzles move to wide teams partway through (both at once, the most
    the serving loop widens) and still return the CPU A* move counts; light queries
    served beside them are unaffected."""
    ctx = sctx
    rows = sorted(json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))), key=lambda r: -r["expanded"])
    heavy = rows[:2]
    light = [r for r in rows if r["expanded"] < 100_000][:4]
    w0 = ctx.serve_stats()["widened"]
    tickets = {ctx.serve_submit(ctx.puzzle_query(4, r["board"])): r["steps"] for r in heavy}
    for r in light:
        tickets[ctx.serve_submit(ctx.puzzle_query(4, r["board"]))] = r["steps"]
    _say(f"heavy puzzles seeds {[r['seed'] for r in heavy]} ({[r['expanded'] for r in heavy]} CPU expansions) "
         f"+ {len(light)} light")
    got = _drain(ctx, len(tickets))
    for t, steps in tickets.items():
        assert got[t]["found"] and got[t]["steps"] == steps, (t, got[t], steps)
    assert ctx.serve_stats()["widened"] > w0
