"""Serving the C5 stream's own queries against the oracle (tests/golden/c5_stream.json:
the first 40 queries of the bench's serving stream, seed 5, plus its three deepest grids,
with answers from the oracle's dense exact A* and the c3 puzzle fixtures). The grids
have sides 588..7786, 13 of them at least 2^22 cells (the large-grid class); the deep
ones run for tens of time-budgeted launches, so they are suspended and resumed many
times. Every cost must equal the fixture's, every served path must be valid on the
query's own bitmap and have exactly that cost, and every move list must replay to the
goal (reference semantics: pathway/CPU-solver.cpp:46-90, puzzle/CPU-solver.hpp:154-260)."""
import json
import os

import pytest

import uastar_amd as U
from tests.helpers import check_grid_path_bitmap, goal_board, replay_moves

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def test_serve_c5_stream_matches_fixtures():
    rows = json.load(open(os.path.join(GOLDEN, "c5_stream.json")))
    with U.Context(0, max_rounds=1_000_000, mem_budget=48 << 30, serve_paths=True) as ctx:
        ctx.puzzle_build_pdb(4)
        grids, queries = {}, []
        for i, r in enumerate(rows):
            if r["kind"] == "grid":
                grids[i] = ctx.grid_generate(r["H"], r["W"], r["seed"], r["p"], *r["start"], *r["goal"])
                queries.append(ctx.grid_query(grids[i], *r["start"], *r["goal"]))
            else:
                queries.append(ctx.puzzle_query(4, r["board"]))
        budget_ms = ctx.serve_stats()["budget_us"] / 1e3
        ticket_of, done = {}, {}
        # staggered arrivals: six queries join before each step, so later ones are admitted
        # while the deep grids are in flight
        pos = 0
        for _ in range(200000):
            for _ in range(6):
                if pos < len(queries):
                    ticket_of[ctx.serve_submit(queries[pos])] = pos
                    pos += 1
            active, pending = ctx.serve_step()
            for t, r in ctx.serve_poll():
                done[ticket_of[t]] = (t, r)
            if pos == len(queries) and active == 0 and pending == 0:
                break
        assert len(done) == len(rows)
        _say(f"served {len(rows)} queries in {ctx.serve_stats()['launches']} launches")
        for i, row in enumerate(rows):
            t, r = done[i]
            tag = f"query {row['index']} ({row['kind']} {row.get('H', '')})"
            assert r["status"] >= 0, (tag, r)
            if row["kind"] == "grid":
                assert bool(r["found"]) == row["found"], (tag, r)
                if not row["found"]:
                    continue
                assert (r["orth"], r["diag"]) == (row["orth"], row["diag"]), (tag, r, row)
                check_grid_path_bitmap(ctx.serve_path(t), grids[i].download(), row["H"], row["W"], *row["start"],
                                       *row["goal"], r["orth"], r["diag"])
                if row["deep"]:  # resident over many launches: suspended and resumed
                    assert r["kernel_ms"] >= 5 * budget_ms, (tag, r)
            else:
                assert r["found"] and r["steps"] == row["steps"], (tag, r, row)
                mv = ctx.serve_path(t).tolist()
                assert len(mv) == row["steps"] and replay_moves(4, row["board"], mv) == goal_board(4), tag
        for g in grids.values():
            g.free()
