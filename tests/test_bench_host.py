"""Host-side logic of bench.py, without a GPU: the one-GPU guard, the serving stream's
generator against the stored stream fixture (so the GPU test of that fixture covers the
stream the bench serves), the source check of the traffic figure, and the FIFO latency
of the plain pass."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_stream_generator_matches_the_stream_fixture():
    """bench.mixed_stream (seed 5) against tests/golden/c5_stream.json, which stores
    queries of the same stream by index with the dense oracle's answers."""
    with open(os.path.join(ROOT, "tests", "golden", "c5_stream.json")) as f:
        fixture = json.load(f)
    specs = bench.mixed_stream(1 + max(q["index"] for q in fixture))
    for q in fixture:
        s = specs[q["index"]]
        assert s["kind"] == q["kind"], q["index"]
        if q["kind"] == "grid":
            assert (s["H"], s["W"], s["seed"], s["p"]) == (q["H"], q["W"], q["seed"], q["p"])
            assert list(s["s"]) == q["start"] and list(s["g"]) == q["goal"]
        else:
            assert s["board"] == q["board"] and s["steps"] == q["steps"]
    # the stated mix: 70 % grids with sides in [512, 8192], 30 % 15-puzzles
    many = bench.mixed_stream(2000, seed=7)
    grids = [s for s in many if s["kind"] == "grid"]
    assert 0.66 < len(grids) / len(many) < 0.74
    assert min(s["H"] for s in grids) >= 512 and max(s["H"] for s in grids) <= 8192
    assert all(0.2 <= s["p"] <= 0.4 for s in grids)


def test_traffic_figure_only_for_its_own_invocation():
    v, src = bench.measured_traffic("--steps 20 --warmup 5 --load 20")
    assert v is not None and v > 1e9 and "profiles" in src
    v, src = bench.measured_traffic("--steps 5 --warmup 1 --load 20")
    assert v is None and "not of this run" in src


def test_fifo_latency_of_the_plain_pass():
    # arrivals 0, 1, 10 ms with service 5, 5, 1 ms: the second waits for the first
    # (done at 5, then 10), the third arrives to an idle server
    assert bench.fifo_latency([0.0, 1.0, 10.0], [5.0, 5.0, 1.0]) == [5.0, 9.0, 1.0]


def test_move_replay_against_the_fixture_solutions():
    """bench.replay_board (the check of the bounded workload's answers) takes every stored
    24-puzzle solution to the goal, and a solution with one move dropped does not."""
    with open(os.path.join(ROOT, "tests", "golden", "c5_puzzles.json")) as f:
        rows = json.load(f)
    goal = list(range(1, 25)) + [0]
    for r in rows:
        assert len(r["moves"]) == r["steps"]
        assert bench.replay_board(5, r["board"], r["moves"]) == goal, r["seed"]
        if r["moves"]:
            assert bench.replay_board(5, r["board"], r["moves"][:-1]) != goal
