"""GPU tests of the 5x5 puzzle (the 24-puzzle, uAstar's N = 5): the four 6-tile pattern
databases built on the device equal the oracle's byte for byte (sha256 fixtures), and
plain, batched and served solves give the optimal steps of the golden boards (IDA*
with Manhattan distance and A* with the oracle's PDB agree on each), with moves that
replay to the goal. 5x5 queries run in a kernel of their own (the SET_FIVE set)."""
import hashlib
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import goal_board, grid_bytes, random_grid, replay_moves

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def _rows(max_expanded):
    rows = json.load(open(os.path.join(GOLDEN, "c5_puzzles.json")))
    return [r for r in rows if r["expanded"] <= max_expanded]


@pytest.fixture(scope="module")
def c5():
    c = U.Context(0, max_rounds=200_000, mem_budget=16 << 30)
    _say("building the 5x5 PDB on the device")
    c.puzzle_build_pdb(5)
    yield c
    c.close()


def test_pdb5_matches_oracle_digests(c5):
    pdb = c5.puzzle_download_pdb(5)
    digests = json.load(open(os.path.join(GOLDEN, "pdb_digests.json")))["5"]
    off = 0
    for d in digests:
        tab = pdb[off: off + d["size"]]
        off += d["size"]
        assert np.bincount(tab, minlength=256).tolist() == d["histogram"], d["tiles"]
        assert hashlib.sha256(tab.tobytes()).hexdigest() == d["sha256"], d["tiles"]
    assert off == pdb.size == U.PDB_BYTES[5]


def test_plain_solves5(c5):
    for r in _rows(3_000_000):
        _say(f"plain 5x5 seed {r['seed']} ({r['steps']} steps)")
        res = c5.puzzle_solve(5, np.array(r["board"], np.uint8))
        assert res["status"] == 0 and res["found"] and res["steps"] == r["steps"], (r["seed"], res)
        mv = c5.puzzle_moves()
        assert len(mv) == r["steps"]
        assert replay_moves(5, r["board"], mv) == goal_board(5)
    res = c5.puzzle_solve(5, np.array(goal_board(5), np.uint8))
    assert res["found"] and res["steps"] == 0
    bad = goal_board(5)
    bad[0], bad[1] = bad[1], bad[0]  # one transposition: unsolvable
    assert c5.puzzle_solve(5, np.array(bad, np.uint8))["status"] == U.UA_NO_PATH


def test_batch_and_serve_mix_5x5(c5, oracle):
    """Grids, 15-puzzles and 24-puzzles in one batch (a launch per problem set) and in
    one serving stream: every answer exact."""
    rng = np.random.default_rng(55)
    grids = []
    for _ in range(6):
        H, W = int(rng.integers(30, 200)), int(rng.integers(30, 200))
        m = random_grid(rng, H, W, 0.3)
        m[0, 0] = True
        g = grid_bytes(m)
        grids.append((c5.grid_from_bytes(g, H, W), g, H, W, oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")))
    p4 = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))) if r["expanded"] < 300_000][:3]
    p5 = _rows(1_500_000)
    qs = [c5.grid_query(h, 0, 0, H - 1, W - 1) for h, _, H, W, _ in grids]
    qs += [c5.puzzle_query(4, np.array(r["board"], np.uint8)) for r in p4]
    qs += [c5.puzzle_query(5, np.array(r["board"], np.uint8)) for r in p5]

    def check(res):
        for (_, _, _, _, e), r in zip(grids, res):
            assert bool(r["found"]) == e["found"] and (not e["found"] or (r["orth"], r["diag"]) == (e["orth"], e["diag"]))
        for row, r in zip(p4 + p5, res[len(grids):]):
            assert r["status"] == 0 and r["found"] and r["steps"] == row["steps"], (row["seed"], r)

    _say(f"batch: {len(grids)} grids, {len(p4)} 15-puzzles, {len(p5)} 24-puzzles")
    before = c5.launch_stats()
    res = c5.batch_solve(qs)
    check(res)
    after = c5.launch_stats()
    assert after["direct_launches"] + after["graph_launches"] - before["direct_launches"] - before["graph_launches"] >= 2
    for i in range(len(grids) + len(p4), len(qs)):  # moves of the batched 24-puzzles
        mv = c5.batch_path(i)
        row = p5[i - len(grids) - len(p4)]
        assert replay_moves(5, row["board"], mv.tolist()) == goal_board(5)

    _say("serving the same mix")
    tickets = {c5.serve_submit(q): i for i, q in enumerate(qs)}
    got = {}
    for _ in range(100000):
        a, p = c5.serve_step()
        got.update(c5.serve_poll())
        if len(got) == len(qs) and a == 0 and p == 0:
            break
    res = [None] * len(qs)
    for t, r in got.items():
        res[tickets[t]] = r
    check(res)
    for h, *_ in grids:
        h.free()


def test_memory_bounded_scheme(c5):
    """The memory-bounded scheme (README.md:30-34 of the reference, never built there): a
    board that plain A* solves in ample memory comes back optimal; in a bound too small
    for plain A* the weighted, pruning search returns a valid solution, flagged as not
    proven optimal, no shorter than the optimum and of the same parity."""
    hard = max(_rows(10_000_000), key=lambda r: r["expanded"])
    board = np.array(hard["board"], np.uint8)
    _say(f"bounded 5x5 seed {hard['seed']} in 4 GiB")
    r = c5.puzzle_solve_bounded(5, board, 4 << 30)
    assert r["status"] == 0 and r["optimal"] == 1 and r["weight"] == 1 and r["steps"] == hard["steps"]
    for mem in (256 << 20, 64 << 20):
        _say(f"bounded 5x5 seed {hard['seed']} in {mem >> 20} MiB")
        r = c5.puzzle_solve_bounded(5, board, mem)
        assert r["status"] == 0 and r["found"], r
        assert r["optimal"] == 0 and r["weight"] >= 2
        assert r["steps"] >= hard["steps"] and (r["steps"] - hard["steps"]) % 2 == 0
        mv = c5.puzzle_moves()
        assert len(mv) == r["steps"] and replay_moves(5, hard["board"], mv) == goal_board(5)
    # a 15-puzzle too: plain A* fits a generous bound
    row = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))) if r["expanded"] < 300_000][0]
    r = c5.puzzle_solve_bounded(4, np.array(row["board"], np.uint8), 1 << 30)
    assert r["status"] == 0 and r["optimal"] == 1 and r["steps"] == row["steps"]
    # and the hardest fixture 15-puzzle in 32 MiB: the plan (16-byte hash slots, heaps)
    # keeps to the bound, and the weighted search returns a valid, longer-or-equal solution
    hard4 = max(json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))), key=lambda r: r["expanded"])
    _say(f"bounded 15-puzzle seed {hard4['seed']} in 32 MiB")
    r = c5.puzzle_solve_bounded(4, np.array(hard4["board"], np.uint8), 32 << 20)
    assert r["status"] == 0 and r["found"], r
    assert r["steps"] >= hard4["steps"] and (r["steps"] - hard4["steps"]) % 2 == 0
    mv = c5.puzzle_moves()
    assert len(mv) == r["steps"] and replay_moves(4, hard4["board"], mv) == goal_board(4)
