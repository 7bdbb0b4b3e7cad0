"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own code: closed forms, a textbook
Dijkstra written here in Python with floating costs, exhaustive BFS of the
8-puzzle, the lexicographic rank of k-permutations from itertools, and IDA*
with a different heuristic.
"""
import heapq
import itertools
import math

import numpy as np
import pytest

from tests.helpers import DX, DY, check_grid_path, goal_board, grid_bytes, replay_moves

SQ2 = math.sqrt(2.0)


def py_dijkstra(graph, H, W, sx, sy, ex, ey):
    """Textbook Dijkstra over (cost as float, orth, diag); float is exact enough to
    separate distinct a + b*sqrt2 on these tiny grids (gaps > 1e-4)."""
    g = np.asarray(graph).reshape(-1)
    dist = {}
    pq = [(0.0, 0, 0, sx * W + sy)]
    while pq:
        c, a, b, u = heapq.heappop(pq)
        if u in dist:
            continue
        dist[u] = (a, b)
        if u == ex * W + ey:
            return True, a, b
        x, y = divmod(u, W)
        for i in range(8):
            if not g[u] & (1 << i):
                continue
            nx, ny = x + DX[i], y + DY[i]
            if 0 <= nx < H and 0 <= ny < W and nx * W + ny not in dist:
                na, nb = a + (i < 4), b + (i >= 4)
                heapq.heappush(pq, (na + nb * SQ2, na, nb, nx * W + ny))
    return False, 0, 0


def octile_pair(sx, sy, ex, ey):
    dx, dy = abs(sx - ex), abs(sy - ey)
    return abs(dx - dy), min(dx, dy)


def test_readme_example(oracle):
    # README.md:13-22 / main.cpp:19-26; hand-derived optimum 4 + 3*sqrt(2)
    m = np.array([[1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 1, 1, 0, 1], [1, 1, 1, 0, 1]])
    g = grid_bytes(m)
    for v in ["ref", "dense", "dijkstra"]:
        r = oracle.grid_astar(g, 5, 5, 0, 0, 4, 4, v)
        assert (r["found"], r["orth"], r["diag"]) == (True, 4, 3)
        check_grid_path(r["path"], g, 5, 5, 0, 0, 4, 4, 4, 3)


@pytest.mark.parametrize("H,W,sx,sy,ex,ey", [(9, 13, 0, 0, 8, 12), (20, 3, 19, 2, 0, 0), (1, 17, 0, 16, 0, 0),
                                             (31, 31, 15, 15, 0, 30)])
def test_empty_grid_is_octile(oracle, H, W, sx, sy, ex, ey):
    g = grid_bytes(np.ones((H, W)))
    r = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
    assert (r["orth"], r["diag"]) == octile_pair(sx, sy, ex, ey)


def zigzag_closed_form(H, W, G):
    """Octile distances summed over the gap waypoints of pathway/input/zigzag.cpp:14-40."""
    pts = [(0, W // 2)]
    k = 1
    while k * G <= H - 1:
        if k * G == H - 1:
            break
        pts.append((k * G, W - 1 if k % 2 == 1 else 0))
        k += 1
    pts.append((H - 1, W // 2))
    a = b = 0
    for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
        da, db = octile_pair(x0, y0, x1, y1)
        a += da
        b += db
    return a, b


@pytest.mark.parametrize("H,W,G", [(10, 7, 3), (21, 11, 2), (15, 5, 7), (2, 9, 5), (50, 40, 6), (12, 4, 11)])
def test_zigzag_closed_form(oracle, H, W, G):
    g, (sx, sy, ex, ey) = oracle.grid_zigzag(H, W, G)
    r = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
    assert r["found"]
    assert (r["orth"], r["diag"]) == zigzag_closed_form(H, W, G)
    check_grid_path(r["path"], g, H, W, sx, sy, ex, ey, r["orth"], r["diag"])


@pytest.mark.parametrize("seed", range(40))
def test_tiny_grids_vs_python_dijkstra(oracle, seed):
    rng = np.random.default_rng(seed)
    H, W = int(rng.integers(1, 14)), int(rng.integers(1, 14))
    m = rng.random((H, W)) >= rng.uniform(0, 0.5)
    sx, sy, ex, ey = (int(rng.integers(H)), int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W)))
    g = grid_bytes(m)
    exp = py_dijkstra(g, H, W, sx, sy, ex, ey)
    for v in ["ref", "dense", "dijkstra"]:
        r = oracle.grid_astar(g, H, W, sx, sy, ex, ey, v)
        assert r["found"] == exp[0]
        if exp[0]:
            assert (r["orth"], r["diag"]) == exp[1:]
            check_grid_path(r["path"], g, H, W, sx, sy, ex, ey, r["orth"], r["diag"])


def test_blocked_goal_reachable_and_blocked_start(oracle):
    m = np.ones((6, 6))
    m[5, 5] = 0
    r = oracle.grid_astar(grid_bytes(m), 6, 6, 0, 0, 5, 5, "ref")
    assert (r["found"], r["orth"], r["diag"]) == (True, 0, 5)
    m = np.ones((6, 6))
    m[0, 0] = 0
    assert not oracle.grid_astar(grid_bytes(m), 6, 6, 0, 0, 5, 5, "ref")["found"]


def test_fp32_variant_within_reference_tolerance(oracle):
    H = W = 300
    g = oracle.grid_bernoulli(H, W, 11, 0.3, 0, 0, H - 1, W - 1)
    ex = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    f = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "f32")
    exact = ex["orth"] + ex["diag"] * SQ2
    # utils.hpp:86-111 float_equal: 1e-3 absolute or 200 ULPs
    assert abs(f["cost_f32"] - exact) <= max(1e-3, 200 * np.spacing(np.float32(exact)))


def test_generators(oracle):
    g = oracle.grid_bernoulli(200, 300, 5, 0.3, 0, 0, 199, 299)
    assert g[0] == 0xFF and g[-1] == 0xFF
    frac = (g == 0).mean()
    assert abs(frac - 0.3) < 0.01
    for s in range(50):
        t = oracle.puzzle_random(4, 3000 + s)
        assert sorted(t.tolist()) == list(range(16))
        assert oracle.puzzle_solvable(4, t)
    # solvability rule vs a reachable / unreachable pair
    goal = np.array(goal_board(4), np.uint8)
    assert oracle.puzzle_solvable(4, goal)
    bad = goal.copy()
    bad[0], bad[1] = bad[1], bad[0]
    assert not oracle.puzzle_solvable(4, bad)


def test_rank_is_lexicographic_kperm(oracle):
    # database.cpp:140-199: the code of a pattern is the lexicographic rank of the
    # tuple of its tiles' cells among k-permutations of the n*n cells
    n, tracked = 3, [2, 5, 7]
    for idx, cells in enumerate(itertools.permutations(range(9), 3)):
        conf = np.zeros(9, np.uint8)
        for t, c in zip(tracked, cells):
            conf[c] = t
        assert oracle.pdb_encode(n, tracked, conf) == idx
        assert np.array_equal(oracle.pdb_decode(n, tracked, idx), conf)
    assert oracle.pdb_size(4, 8) == 518918400 and oracle.pdb_size(4, 7) == 57657600


def eight_puzzle_bfs():
    """Exhaustive BFS of the 8-puzzle from the goal (blank moves)."""
    start = tuple(goal_board(3))
    dist = {start: 0}
    frontier = [start]
    while frontier:
        nxt = []
        for s in frontier:
            z = s.index(0)
            x, y = divmod(z, 3)
            for i in range(4):
                nx, ny = x + DX[i], y + DY[i]
                if 0 <= nx < 3 and 0 <= ny < 3:
                    b = list(s)
                    b[z], b[nx * 3 + ny] = b[nx * 3 + ny], b[z]
                    t = tuple(b)
                    if t not in dist:
                        dist[t] = dist[s] + 1
                        nxt.append(t)
        frontier = nxt
    return dist


@pytest.fixture(scope="module")
def bfs8():
    return eight_puzzle_bfs()


def test_pdb3_equals_exhaustive_bfs(oracle, bfs8):
    assert len(bfs8) == 181440 and max(bfs8.values()) == 31
    assert sum(1 for v in bfs8.values() if v == 31) == 2
    tracked = list(range(1, 9))
    for method in (0, 1):
        p = oracle.pdb_build(3, tracked, method)
        assert int((p != 255).sum()) == 181440
        for s, d in list(bfs8.items())[::97]:
            assert p[oracle.pdb_encode(3, tracked, np.array(s, np.uint8))] == d


def test_solvable_rule_is_reachability_over_all_eight_puzzle_boards(oracle, bfs8):
    """orc_puzzle_solvable (the parity rule of the generators) against the exhaustive BFS:
    over all 9! boards it holds exactly for the 181,440 reachable from the goal."""
    import itertools
    n_true = 0
    for perm in itertools.permutations(range(9)):
        ok = oracle.puzzle_solvable(3, np.array(perm, np.uint8))
        assert ok == (perm in bfs8), perm
        n_true += ok
    assert n_true == 181440


def test_eight_puzzle_astar_vs_bfs(oracle, bfs8):
    pdb = oracle.pdb_full(3)
    rng = np.random.default_rng(3)
    states = list(bfs8.keys())
    for i in rng.choice(len(states), 60, replace=False):
        s = np.array(states[i], np.uint8)
        steps, moves, _ = oracle.puzzle_solve(3, s, pdb, "astar")
        assert steps == bfs8[states[i]]
        assert replay_moves(3, s, moves) == goal_board(3)
        assert len(moves) == steps


def test_small_pattern_pdb_queue_equals_levels(oracle):
    for n, tracked in ((4, [1, 2, 3]), (4, [5, 9, 13, 14]), (4, [15]), (5, [2, 1, 6]), (5, [13, 14, 18, 19])):
        a = oracle.pdb_build(n, tracked, 0)
        b = oracle.pdb_build(n, tracked, 1)
        assert np.array_equal(a, b)
        assert not (a == 255).any()   # >= 8 free cells: every placement is reachable


def random_walk_board(rng, n, steps):
    b = goal_board(n)
    z = b.index(0)
    last = -1
    for _ in range(steps):
        opts = [i for i in range(4) if 0 <= z // n + DX[i] < n and 0 <= z % n + DY[i] < n and i != (last ^ 1)]
        i = int(rng.choice(opts))
        nz = (z // n + DX[i]) * n + z % n + DY[i]
        b[z], b[nz] = b[nz], b[z]
        z, last = nz, i
    return np.array(b, np.uint8)


@pytest.mark.slow
def test_fifteen_puzzle_pdb_and_astar(oracle):
    """The 7-8 PDB (built with the level BFS) against its invariants, and A* with it
    against IDA* with Manhattan distance on boards a random walk away from the goal."""
    pdb = oracle.pdb_full(4, 1, cache_dir="_cache")
    assert pdb.size == 576576000
    parts = oracle.partitions(4)
    assert parts == [[1, 2, 3, 4, 5, 6, 9, 13], [7, 8, 10, 11, 12, 14, 15]]
    assert not (pdb == 255).any()
    goal = np.array(goal_board(4), np.uint8)
    assert oracle.puzzle_h(4, goal, pdb) == 0
    rng = np.random.default_rng(5)
    for k in range(12):
        b = random_walk_board(rng, 4, int(rng.integers(10, 40)))
        s1, m1, _ = oracle.puzzle_solve(4, b, pdb, "astar")
        s2, _, _ = oracle.puzzle_solve(4, b, None, "ida_manhattan", node_limit=50_000_000)
        assert s1 == s2
        assert replay_moves(4, b, m1) == goal_board(4)
        h = oracle.puzzle_h(4, b, pdb)
        assert h <= s1 and (s1 - h) % 2 == 0
    # consistency: neighbouring boards differ in h by at most 1
    for _ in range(200):
        b = oracle.puzzle_random(4, int(rng.integers(1 << 40)))
        z = int(np.flatnonzero(b == 0)[0])
        h0 = oracle.puzzle_h(4, b, pdb)
        for i in range(4):
            x, y = z // 4 + DX[i], z % 4 + DY[i]
            if 0 <= x < 4 and 0 <= y < 4:
                c = b.copy()
                c[z], c[x * 4 + y] = c[x * 4 + y], c[z]
                assert abs(oracle.puzzle_h(4, c, pdb) - h0) == 1


@pytest.mark.slow
def test_twentyfour_puzzle_pdb_and_astar(oracle):
    """The reference's 6-6-6-6 PDB of the 5x5 board (level BFS) against its invariants,
    and A* with it against IDA* with Manhattan distance on walks off the goal."""
    parts = oracle.partitions(5)
    assert parts == [[3, 4, 5, 10, 15, 20], [2, 1, 6, 11, 16, 21], [7, 8, 9, 12, 17, 22], [13, 14, 18, 19, 23, 24]]
    pdb = oracle.pdb_full(5, 1, cache_dir="_cache")
    assert pdb.size == 4 * 127512000
    assert not (pdb == 255).any()
    goal = np.array(goal_board(5), np.uint8)
    assert oracle.puzzle_h(5, goal, pdb) == 0
    rng = np.random.default_rng(25)
    for _ in range(8):
        b = random_walk_board(rng, 5, int(rng.integers(10, 36)))
        s1, m1, _ = oracle.puzzle_solve(5, b, pdb, "astar")
        s2, _, _ = oracle.puzzle_solve(5, b, None, "ida_manhattan", node_limit=50_000_000)
        assert s1 == s2
        assert replay_moves(5, b, m1) == goal_board(5)
        h = oracle.puzzle_h(5, b, pdb)
        assert h <= s1 and (s1 - h) % 2 == 0
    for _ in range(200):  # consistency: a blank move changes h by exactly one
        b = oracle.puzzle_random(5, int(rng.integers(1 << 40)))
        z = int(np.flatnonzero(b == 0)[0])
        h0 = oracle.puzzle_h(5, b, pdb)
        for i in range(4):
            x, y = z // 5 + DX[i], z % 5 + DY[i]
            if 0 <= x < 5 and 0 <= y < 5:
                c = b.copy()
                c[z], c[x * 5 + y] = c[x * 5 + y], c[z]
                assert abs(oracle.puzzle_h(5, c, pdb) - h0) == 1
