"""The uastar command line (uAstar's CLI and input modules on libuastar): CPU-side runs
checked against the oracle and the golden PDB digests, and (GPU) the CPU / GPU
consistency check of a full run."""
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

from tests.helpers import goal_board

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "uastar_amd", "uastar")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(args, stdin="", cwd=None, timeout=300):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is missing: build it with `python -m uastar_amd.build`")
    p = subprocess.run([CLI, *args], input=stdin, capture_output=True, text=True, cwd=cwd, timeout=timeout)
    return p.returncode, p.stdout


def costs(out, who):
    """The exact (orth, diag) pair of the run, checked against the reference's own line
    ' > Optimal distance from WHO: %.3f' (pathway/pathway.cpp:115-119), which must be
    present, exactly so formatted, and agree with the pair."""
    m = re.search(rf"^ > Exact distance from {who}: (\d+) \+ (\d+) \* sqrt\(2\)$", out, re.M)
    if not m:
        assert not re.search(rf"Optimal distance from {who}", out), out
        return None
    pair = (int(m.group(1)), int(m.group(2)))
    line = re.search(rf"^ > Optimal distance from {who}: (\d+\.\d{{3}})$", out, re.M)
    assert line, out
    assert line.group(1) == f"{pair[0] + pair[1] * 2 ** 0.5:.3f}"
    return pair


def test_readme_example(tmp_path):
    grid = "0 0 4 4\n1 0 1 1 1\n1 0 1 1 1\n1 0 1 1 1\n1 1 1 0 1\n1 1 1 0 1\n"
    rc, out = run(["--pathway", "-H", "5", "-W", "5", "--input-module", "custom", "--no-gpu"], grid, cwd=tmp_path)
    assert rc == 0, out
    assert costs(out, "CPU") == (4, 3)
    assert " > Optimal distance from CPU: 8.243\n" in out
    # the lines of Pathway::printSolution and CPUPathwaySolver::getSolution
    assert " > Number of -: 4, +: 3\n" in out
    assert re.search(r"^\t\t\tNumber of nodes expanded: \d+$", out, re.M)
    cells = re.findall(r"\((\d+) (\d+)\)", (tmp_path / "solutionCPU.txt").read_text())
    assert cells[0] == ("0", "0") and cells[-1] == ("4", "4") and len(cells) == 8


@pytest.mark.parametrize("H,W,gap", [(10, 7, 3), (31, 12, 5), (2, 9, 1)])
def test_zigzag_matches_oracle(oracle, tmp_path, H, W, gap):
    g, (sx, sy, ex, ey) = oracle.grid_zigzag(H, W, gap)
    ref = oracle.grid_astar(g, H, W, sx, sy, ex, ey, "ref")
    rc, out = run(["--pathway", "-H", str(H), "-W", str(W), "--input-module", "zigzag", "-C"], f"{gap}\n", cwd=tmp_path)
    assert rc == 0, out
    if ref["found"]:
        assert costs(out, "CPU") == (ref["orth"], ref["diag"])
    else:
        assert "No solution from CPU." in out


@pytest.mark.parametrize("seed,rate", [(1, 30), (7, 40)])
def test_random_module_matches_oracle(oracle, tmp_path, seed, rate):
    H, W = 120, 90
    g = oracle.grid_bernoulli(H, W, seed, rate / 100.0, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    rc, out = run(["--pathway", "-H", str(H), "-W", str(W), "--input-module", "random", "-b", str(rate),
                   "-s", str(seed), "-C", "-p", "g.bmp"], cwd=tmp_path)
    assert rc == 0, out
    assert costs(out, "CPU") == ((ref["orth"], ref["diag"]) if ref["found"] else None)
    assert (tmp_path / "g.bmp").read_bytes()[:2] == b"BM"


def test_puzzle3_cache_file_and_solution(oracle, tmp_path):
    board = [8, 1, 3, 4, 0, 2, 7, 6, 5]
    steps, _, _ = oracle.puzzle_solve(3, np.array(board, np.uint8), oracle.pdb_full(3), "astar")
    rc, out = run(["--puzzle", "-W", "3", "-C", "--pdb-dir", str(tmp_path)], " ".join(map(str, board)), cwd=tmp_path)
    assert rc == 0, out
    assert f"\n > Optimal steps from CPU: {steps}\n" in out
    assert re.search(r"^\t\tNumber of nodes deduplicated: \d+\n\t\tNumber of nodes expanded: \d+$", out, re.M)
    digest = json.load(open(os.path.join(GOLDEN, "pdb_digests.json")))["3"][0]
    name = "database_3_" + "_".join(map(str, digest["tiles"])) + ".bin"
    assert hashlib.sha256((tmp_path / name).read_bytes()).hexdigest() == digest["sha256"]
    states = [b for b in (tmp_path / "solutionCPU.txt").read_text().split("\n\n") if b.strip()]
    assert len(states) == steps + 1
    assert [int(v) for v in states[-1].split()] == goal_board(3)
    # a second run reads the cache file instead of building it
    rc, out = run(["--puzzle", "-H", "3", "-C", "--pdb-dir", str(tmp_path)], " ".join(map(str, board)), cwd=tmp_path)
    assert rc == 0 and "Generating pattern database" not in out


def test_rejects_bad_input(tmp_path):
    rc, out = run([], cwd=tmp_path)
    assert rc == 1 and "--pathway" in out
    rc, out = run(["--pathway", "--puzzle", "-H", "3"], cwd=tmp_path)
    assert rc == 1 and "select the type of problem" in out
    rc, out = run(["--pathway", "-H", "5"], cwd=tmp_path)
    assert rc == 1 and "width and height" in out
    rc, out = run(["--puzzle", "-W", "3", "-C"], "1 2 3 4 5 6 7 8 8", cwd=tmp_path)
    assert rc == 1 and "each of 0..8 once" in out
    rc, out = run(["--puzzle", "-W", "5", "-C"], cwd=tmp_path)
    assert rc == 1
    rc, out = run(["--pathway", "-H", "4", "-W", "4", "--input-module", "nope", "-C"], cwd=tmp_path)
    assert rc == 1 and "input-module" in out


def test_unsolvable_puzzle_reports_no_solution(tmp_path):
    rc, out = run(["--puzzle", "-W", "3", "-C", "--pdb-dir", str(tmp_path)], "2 1 3 4 5 6 7 8 0", cwd=tmp_path)
    assert rc == 0 and "No solution from CPU." in out


@pytest.mark.gpu
def test_cli_cpu_and_gpu_agree(tmp_path):
    """Full runs (CPU and GPU, uAstar's consistency check) on a random grid, the README
    grid and an 8-puzzle; a 15-puzzle and a 24-puzzle on the GPU alone against the
    fixtures. (The CLI runs while this test process holds no device context.)"""
    print("cli: random grid 400x300", flush=True)
    rc, out = run(["--pathway", "-H", "400", "-W", "300", "--input-module", "random", "-b", "30", "-s", "5"],
                  cwd=tmp_path)
    assert rc == 0, out
    assert costs(out, "CPU") == costs(out, "GPU") and costs(out, "GPU") is not None
    assert re.search(r"^\t\t\t Number of nodes expanded: \d+$", out, re.M), out
    gpu_cells = re.findall(r"\((\d+) (\d+)\)", (tmp_path / "solutionGPU.txt").read_text())
    assert gpu_cells[0] == ("0", "0") and gpu_cells[-1] == ("399", "299")
    print("cli: 8-puzzle", flush=True)
    rc, out = run(["--puzzle", "-W", "3", "--pdb-dir", str(tmp_path)], "8 1 3 4 0 2 7 6 5", cwd=tmp_path)
    assert rc == 0, out
    assert "Optimal steps from GPU: 14" in out and "Optimal steps from CPU: 14" in out
    row = [r for r in json.load(open(os.path.join(GOLDEN, "c3_puzzles.json"))) if r["expanded"] < 300_000][0]
    print(f"cli: 15-puzzle seed {row['seed']}", flush=True)
    rc, out = run(["--puzzle", "-W", "4", "-G"], " ".join(map(str, row["board"])), cwd=tmp_path)
    assert rc == 0, out
    assert f"Optimal steps from GPU: {row['steps']}" in out
    states = [b for b in (tmp_path / "solutionGPU.txt").read_text().split("\n\n") if b.strip()]
    assert [int(v) for v in states[-1].split()] == goal_board(4)
    assert len(states) == row["steps"] + 1
    row = [r for r in json.load(open(os.path.join(GOLDEN, "c5_puzzles.json"))) if r["expanded"] < 1_000_000][0]
    print(f"cli: 24-puzzle seed {row['seed']}", flush=True)
    rc, out = run(["--puzzle", "-W", "5", "-G"], " ".join(map(str, row["board"])), cwd=tmp_path)
    assert rc == 0, out
    assert f"Optimal steps from GPU: {row['steps']}" in out
    states = [b for b in (tmp_path / "solutionGPU.txt").read_text().split("\n\n") if b.strip()]
    assert [int(v) for v in states[-1].split()] == goal_board(5)
    assert len(states) == row["steps"] + 1
