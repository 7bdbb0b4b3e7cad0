"""The uAstar-side binding (uastar_amd/integration/): GPUPathwaySolver and
GPUPuzzleSolver<N> implemented over libuastar's C ABI, compiled against declarations of
uAstar's own classes (tests/binding/, after src/pathway/GPU-solver.hpp:19-35,
src/pathway/pathway.hpp:10-33, src/puzzle/puzzle.cuh:6-38, src/problem.hpp:6-24) and
linked with libuastar.so. On the CPU: it builds and links, and without a GPU it fails
loudly with libuastar's status. On the GPU: it solves through the classes' own calls,
checked against the oracle, and reuses uAstar's PDB cache files. (This module runs
before any test of this process holds a device context, so the binding's own process
can make cooperative launches.)"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path, goal_board, replay_moves

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def binding(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("binding") / "binding_main")
    libdir = os.path.join(ROOT, "uastar_amd")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "tests", "binding"),
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "uastar_amd", "integration"),
           os.path.join(ROOT, "tests", "binding", "binding_main.cpp"),
           os.path.join(ROOT, "uastar_amd", "integration", "pathway_gpu_solver.cpp"),
           "-L", libdir, "-luastar", f"-Wl,-rpath,{libdir}", "-o", exe]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def run(exe, args, cwd, timeout=300):
    p = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=timeout)
    return p.returncode, p.stdout


def test_binding_compiles_links_and_fails_loudly_without_gpu(binding, tmp_path):
    g = np.full(16, 0xFF, np.uint8)
    (tmp_path / "g.bin").write_bytes(g.tobytes())
    if U.device_count() > 0:
        pytest.skip("a GPU is visible: the no-device path is not reachable")
    rc, out = run(binding, ["pathway", 4, 4, 0, 0, 3, 3, "g.bin"], tmp_path)
    assert rc == 2 and "ua_create failed (status -6)" in out, out
    rc, out = run(binding, ["puzzle", 3, 1, 2, 3, 4, 5, 6, 7, 0, 8], tmp_path)
    assert rc == 2 and "ua_create failed (status -6)" in out, out


@pytest.mark.gpu
def test_binding_pathway_matches_oracle(binding, oracle, tmp_path):
    H, W = 80, 60
    print("binding: pathway 80x60", flush=True)
    g = oracle.grid_bernoulli(H, W, 11, 0.3, 0, 0, H - 1, W - 1)
    ref = oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")
    (tmp_path / "g.bin").write_bytes(np.asarray(g, np.uint8).tobytes())
    rc, out = run(binding, ["pathway", H, W, 0, 0, H - 1, W - 1, "g.bin", 0], tmp_path)
    assert rc == 0, out
    assert f"found {int(ref['found'])}" in out
    if ref["found"]:
        assert "\t\t\t Number of nodes expanded: " in out
        cost = float(out.split("cost ")[1].split()[0])
        assert abs(cost - (ref["orth"] + ref["diag"] * 2 ** 0.5)) < 1e-3
        path = [int(v) for v in out.split("path")[1].split()]
        check_grid_path(path, g, H, W, 0, 0, H - 1, W - 1, ref["orth"], ref["diag"])
    # a partial out-move mask cannot be held by the bitmap: an error, not a wrong answer
    bad = np.asarray(g, np.uint8).copy()
    bad[5] = 0x0F
    (tmp_path / "bad.bin").write_bytes(bad.tobytes())
    rc, out = run(binding, ["pathway", H, W, 0, 0, H - 1, W - 1, "bad.bin"], tmp_path)
    assert rc == 2 and "ua_grid_from_bytes failed (status -1)" in out, out


@pytest.mark.gpu
def test_binding_puzzle_builds_then_reuses_cache_files(binding, oracle, tmp_path):
    board = [8, 1, 3, 4, 0, 2, 7, 6, 5]
    steps, _, _ = oracle.puzzle_solve(3, np.array(board, np.uint8), oracle.pdb_full(3), "astar")
    print("binding: 8-puzzle, cache files absent", flush=True)
    rc, out = run(binding, ["puzzle", 3, *board, 0], tmp_path)
    assert rc == 0, out
    assert "Generating pattern databse" in out and f"steps {steps}" in out
    digest = json.load(open(os.path.join(GOLDEN, "pdb_digests.json")))["3"][0]
    name = "database_3_" + "_".join(map(str, digest["tiles"])) + ".bin"
    assert hashlib.sha256((tmp_path / name).read_bytes()).hexdigest() == digest["sha256"]
    moves = [int(v) for v in out.split("moves")[1].split()]
    assert len(moves) == steps and replay_moves(3, board, moves) == goal_board(3)
    print("binding: 8-puzzle, cache files present", flush=True)
    rc, out = run(binding, ["puzzle", 3, *board], tmp_path)
    assert rc == 0, out
    assert "Generating" not in out and f"steps {steps}" in out
    rc, out = run(binding, ["puzzle", 3, 2, 1, 3, 4, 5, 6, 7, 8, 0], tmp_path)  # unsolvable: false at once
    assert rc == 0 and "found 0" in out, out
