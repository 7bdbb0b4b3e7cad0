"""GPU parity of the sliding-tile path: device-built PDBs and optimal move counts.

Bar: PDB tables byte-identical to the oracle's BFS; move counts equal to the oracle's
A*; returned moves replay to the goal with exactly that many moves.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import goal_board, replay_moves

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def test_pdb3_byte_exact(ctx, oracle):
    dev = ctx.puzzle_download_pdb(3)
    ref = oracle.pdb_build(3, list(range(1, 9)), 1)
    assert np.array_equal(dev, ref)


def test_pdb_load_from_host(ctx, oracle):
    """A PDB uploaded from host memory (as from uAstar's cache files) serves the solver."""
    ref = oracle.pdb_build(3, list(range(1, 9)), 1)
    with U.Context(0) as c2:
        with pytest.raises(U.UAError):
            c2.puzzle_load_pdb(3, ref[:-1])  # wrong size
        c2.puzzle_load_pdb(3, ref)
        assert np.array_equal(c2.puzzle_download_pdb(3), ref)
        board = oracle.puzzle_random(3, 777)
        steps, _, _ = oracle.puzzle_solve(3, board, ref, "astar")
        assert c2.puzzle_solve(3, board)["steps"] == steps


def test_pdb4_matches_digest(ctx):
    with open(os.path.join(GOLDEN, "pdb_digests.json")) as f:
        tabs = json.load(f)["4"]
    _say("building the 7-8 PDB on the device")
    dev = ctx.puzzle_download_pdb(4)
    off = 0
    for t in tabs:
        part = dev[off: off + t["size"]]
        off += t["size"]
        assert np.bincount(part, minlength=256).tolist() == t["histogram"]
        assert hashlib.sha256(part.tobytes()).hexdigest() == t["sha256"]


def test_eight_puzzles(ctx, oracle):
    pdb = oracle.pdb_full(3)
    for s in range(24):
        b = oracle.puzzle_random(3, 500 + s)
        exp, _, _ = oracle.puzzle_solve(3, b, pdb, "astar")
        r = ctx.puzzle_solve(3, b)
        mv = ctx.puzzle_moves()
        assert r["found"] and r["steps"] == exp, (s, b.tolist(), r, exp)
        assert len(mv) == exp and replay_moves(3, b, mv) == goal_board(3)


def test_puzzle_edge_cases(ctx):
    goal = np.array(goal_board(4), np.uint8)
    r = ctx.puzzle_solve(4, goal)
    assert r["found"] and r["steps"] == 0 and ctx.puzzle_moves() == []
    bad = goal.copy()
    bad[0], bad[1] = bad[1], bad[0]          # odd permutation: unsolvable
    r = ctx.puzzle_solve(4, bad)
    assert not r["found"] and r["status"] == 1
    one = goal.copy()
    one[14], one[15] = one[15], one[14]      # one move away
    r = ctx.puzzle_solve(4, one)
    assert r["steps"] == 1 and ctx.puzzle_moves() == [2]


def _c3_rows(max_expanded):
    path = os.path.join(GOLDEN, "c3_puzzles.json")
    with open(path) as f:
        rows = json.load(f)
    return [r for r in rows if r["expanded"] <= max_expanded]


def test_fifteen_puzzles_plain(ctx):
    rows = _c3_rows(1_500_000)[:16]
    assert rows
    for r in rows:
        _say(f"15-puzzle seed {r['seed']} (oracle {r['steps']} moves, {r['expanded']} expanded)")
        b = np.array(r["board"], np.uint8)
        res = ctx.puzzle_solve(4, b)
        mv = ctx.puzzle_moves()
        assert res["found"] and res["steps"] == r["steps"], (r["seed"], res)
        assert len(mv) == r["steps"] and replay_moves(4, b, mv) == goal_board(4)
