"""GPU tests of graph-replayed launches. With graphs=True every search launch of a
context -- plain solve, batch sub-batch, serving step -- replays an instantiated
hipGraph (copy of the launch's I/O block in, cooperative kernel between two timing
events, copy of the results out); by default the same operations are issued call by
call. The replays run many times, with the program's own work (table clears,
admission, retirement, synchronisation) between them; each replay's results are
checked against the oracle and against the same queries issued call by call."""
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path, grid_bytes, random_grid

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def _puzzles(k):
    rows = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))) if r["expanded"] < 300_000]
    return rows[:k]


def _grids(seed, k, lo=16, hi=260):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        H, W = int(rng.integers(lo, hi)), int(rng.integers(lo, hi))
        m = random_grid(rng, H, W, float(rng.uniform(0.1, 0.4)))
        sx, sy, ex, ey = int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W))
        m[sx, sy] = True
        out.append((grid_bytes(m), H, W, sx, sy, ex, ey))
    return out


def _key(r):
    return (r["status"], r["found"], r["orth"], r["diag"], r["steps"])


@pytest.fixture(scope="module")
def pair():
    cg = U.Context(0, max_rounds=200_000, mem_budget=8 << 30, graphs=True)
    cd = U.Context(0, max_rounds=200_000, mem_budget=8 << 30)
    yield cg, cd
    cg.close()
    cd.close()


def test_plain_solves_replay_one_graph(pair, oracle):
    cg, cd = pair
    for i, (g, H, W, sx, sy, ex, ey) in enumerate(_grids(21, 8)):
        _say(f"plain grid {i}: {H}x{W}")
        e = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
        got = []
        for ctx in (cg, cd):
            grid = ctx.grid_from_bytes(g, H, W)
            r = ctx.grid_solve(grid, sx, sy, ex, ey)
            assert bool(r["found"]) == bool(e["found"])
            if e["found"]:
                assert (r["orth"], r["diag"]) == (e["orth"], e["diag"])
                check_grid_path(ctx.grid_path(), g, H, W, sx, sy, ex, ey, r["orth"], r["diag"])
            got.append(_key(r))
            grid.free()
        assert got[0] == got[1]
    for row in _puzzles(3):
        _say(f"plain 15-puzzle seed {row['seed']}")
        for ctx in (cg, cd):
            r = ctx.puzzle_solve(4, np.array(row["board"], np.uint8))
            assert r["found"] and r["steps"] == row["steps"]
    sg, sd = cg.launch_stats(), cd.launch_stats()
    assert sg["graph_error"] == "", sg["graph_error"]
    assert sg["graph_launches"] >= 11 and sg["direct_launches"] == 0
    assert sd["graph_launches"] == 0 and sd["direct_launches"] >= 11
    assert sg["wall_ms"] >= sg["kernel_ms"] > 0


def test_batch_replays_match_direct(pair, oracle):
    cg, cd = pair
    grids = _grids(22, 40)
    rows = _puzzles(4)
    res = []
    for ctx in (cg, cd):
        handles = [ctx.grid_from_bytes(g, H, W) for g, H, W, *_ in grids]
        qs = [ctx.grid_query(h, sx, sy, ex, ey) for h, (_, _, _, sx, sy, ex, ey) in zip(handles, grids)]
        qs += [ctx.puzzle_query(4, np.array(r["board"], np.uint8)) for r in rows]
        _say(f"batch of {len(qs)} ({'graph' if ctx is cg else 'direct'})")
        res.append([_key(r) for r in ctx.batch_solve(qs)])
        for h in handles:
            h.free()
    assert res[0] == res[1]
    for (g, H, W, sx, sy, ex, ey), k in zip(grids[:12], res[0]):
        e = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
        assert bool(k[1]) == bool(e["found"]) and (not e["found"] or k[2:4] == (e["orth"], e["diag"]))
    for r, k in zip(rows, res[0][len(grids):]):
        assert k[1] and k[4] == r["steps"]
    assert cg.launch_stats()["direct_launches"] == 0


def test_serving_steps_replay_and_match_direct(pair):
    """Many serving steps of one graph each: queries arrive between steps, finish at
    different steps and suspend in others; the final answers equal the call-by-call ones."""
    cg, cd = pair
    grids = _grids(23, 30, 40, 500)
    rows = _puzzles(5)
    answers = []
    for ctx in (cg, cd):
        before = ctx.launch_stats()
        handles = [ctx.grid_from_bytes(g, H, W) for g, H, W, *_ in grids]
        qs = [ctx.grid_query(h, sx, sy, ex, ey) for h, (_, _, _, sx, sy, ex, ey) in zip(handles, grids)]
        qs += [ctx.puzzle_query(4, np.array(r["board"], np.uint8)) for r in rows]
        tickets, got = {}, {}
        for i, q in enumerate(qs):
            tickets[ctx.serve_submit(q)] = i
            if i % 5 == 4:  # arrivals between steps
                ctx.serve_step()
                got.update(ctx.serve_poll())
        for _ in range(100000):
            a, p = ctx.serve_step()
            got.update(ctx.serve_poll())
            if len(got) == len(qs) and a == 0 and p == 0:
                break
        assert len(got) == len(qs)
        ans = [None] * len(qs)
        for t, r in got.items():
            ans[tickets[t]] = _key(r)
        answers.append(ans)
        after = ctx.launch_stats()
        steps = after["graph_launches"] + after["direct_launches"] - before["graph_launches"] - before["direct_launches"]
        _say(f"served {len(qs)} queries in {steps} launches ({'graph' if ctx is cg else 'direct'})")
        assert steps >= 7
        for h in handles:
            h.free()
    assert answers[0] == answers[1]
    for r, k in zip(rows, answers[0][len(grids):]):
        assert k[1] and k[4] == r["steps"]
    assert cg.launch_stats()["direct_launches"] == 0 and cd.launch_stats()["graph_launches"] == 0
