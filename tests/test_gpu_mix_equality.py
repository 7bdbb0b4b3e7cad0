"""The same 15-puzzles and grid give the same answers alone and in a mix: plain (one
query per launch), in a batch of one, inside a mixed batch with other grids and puzzles
on their own teams, and served alone and in a served mix, each against the fixture / the
oracle. (Boards of the c3 fixture with 0.4 M - 2 M A* expansions, so the puzzles run many
rounds on their multi-block teams beside one-block grid teams on bucket rings.)"""
import json
import os

import pytest

import uastar_amd as U
from tests.helpers import goal_board, replay_moves

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def _drain(ctx, n):
    got = {}
    for _ in range(200000):
        a, p = ctx.serve_step()
        for t, r in ctx.serve_poll():
            got[t] = r
        if len(got) >= n and a == 0 and p == 0:
            break
    return got


def test_puzzles_and_grid_alone_and_in_mix(oracle):
    rows = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json")))
            if 400_000 <= r["expanded"] <= 2_000_000][:4]
    assert len(rows) == 4
    H = W = 350
    g = oracle.grid_bernoulli(H, W, 91, 0.28, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref", want_path=False)
    pair = (ref["orth"], ref["diag"])
    with U.Context(0, max_rounds=1_000_000) as ctx:
        ctx.puzzle_build_pdb(4)
        grid = ctx.grid_generate(H, W, 91, 0.28, 0, 0, H - 1, W - 1)
        fill = [ctx.grid_generate(80 + 10 * k, 90 + 7 * k, 300 + k, 0.3, 0, 0, 79 + 10 * k, 89 + 7 * k)
                for k in range(20)]
        qg = ctx.grid_query(grid, 0, 0, H - 1, W - 1)
        qp = [ctx.puzzle_query(4, r["board"]) for r in rows]
        qf = [ctx.grid_query(f, 0, 0, f.H - 1, f.W - 1) for f in fill]
        _say("plain and batch of one")
        for q, r in zip(qp, rows):
            assert ctx.batch_solve([q])[0]["steps"] == r["steps"]
        for r in rows:
            assert ctx.puzzle_solve(4, r["board"])["steps"] == r["steps"]
        pg = ctx.grid_solve(grid, 0, 0, H - 1, W - 1)
        assert (pg["orth"], pg["diag"]) == pair
        _say("mixed batch")
        mix = [qp[0], qg] + qf[:10] + qp[1:] + qf[10:] + [qg]
        res = ctx.batch_solve(mix)
        steps = [res[0]["steps"]] + [res[12 + k]["steps"] for k in range(3)]
        assert steps == [r["steps"] for r in rows]
        for k, r in enumerate(rows):
            i = 0 if k == 0 else 12 + k - 1
            mv = ctx.batch_path(i).tolist()
            assert len(mv) == r["steps"] and replay_moves(4, r["board"], mv) == goal_board(4)
        assert (res[1]["orth"], res[1]["diag"]) == pair and (res[-1]["orth"], res[-1]["diag"]) == pair
        _say("served alone, then in a served mix")
        alone = {ctx.serve_submit(q): r["steps"] for q, r in zip(qp[:1], rows[:1])}
        done = _drain(ctx, 1)
        assert [done[t]["steps"] for t in alone] == [rows[0]["steps"]]
        tickets = {}
        for k, q in enumerate(qf + [qg] + qp):
            tickets[ctx.serve_submit(q)] = k
        done = _drain(ctx, len(tickets))
        assert len(done) == len(tickets)
        for t, k in tickets.items():
            r = done[t]
            assert r["status"] >= 0, r
            if k == len(qf):
                assert (r["orth"], r["diag"]) == pair
            elif k > len(qf):
                assert r["steps"] == rows[k - len(qf) - 1]["steps"]
        for f in fill:
            f.free()
        grid.free()
