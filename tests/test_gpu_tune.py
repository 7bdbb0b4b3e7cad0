"""GPU tests of the autotuner: it measures every candidate, keeps settings from the
candidate sets, and the tuned context still returns the oracle's answers."""
import json
import os

import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import grid_bytes, random_grid

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_autotune_then_solve(oracle):
    rng = np.random.default_rng(31)
    with U.Context(0, max_rounds=200_000) as ctx:
        grids, queries, expect = [], [], []
        for _ in range(8):
            H, W = int(rng.integers(64, 400)), int(rng.integers(64, 400))
            m = random_grid(rng, H, W, float(rng.uniform(0.1, 0.35)))
            m[0, 0] = m[H - 1, W - 1] = True
            g = grid_bytes(m)
            grid = ctx.grid_from_bytes(g, H, W)
            grids.append(grid)
            queries.append(ctx.grid_query(grid, 0, 0, H - 1, W - 1))
            expect.append(("grid", oracle.grid_astar(g, H, W, 0, 0, H - 1, W - 1, "ref")))
        with open(os.path.join(GOLDEN, "c3_puzzles.json")) as f:
            rows = [r for r in json.load(f) if r["expanded"] < 200_000][:4]
        for r in rows:
            queries.append(ctx.puzzle_query(4, r["board"]))
            expect.append(("puzzle", r["steps"]))
        print(f"autotune on {len(queries)} queries", flush=True)
        t = ctx.autotune(queries)
        print(f"tuning: {t}", flush=True)
        assert t["grid_team_waves"] in (16, 32) and t["grid_large_team_waves"] in (16, 32, 64)
        assert t["puzzle_team_waves"] in (16, 32, 64, 128)
        assert 18 <= t["puzzle_hash_log2"] <= 26 and 18 <= t["plain_hash_log2"] <= 26
        assert t["plain_grid_heaps"] in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)
        assert t["plain_puzzle_heaps"] in (64, 128, 256, 512, 1024, 2048, 4096)
        # every sample grid is below the large-grid threshold: small grids and puzzles are measured
        assert t["trials"] >= 2 + 4 + 2 and t["est_grid_ms"] > 0 and t["est_puzzle_ms"] > 0
        assert t["est_grid_large_ms"] == 0
        got = ctx.get_tuning()
        for k in ("grid_team_waves", "puzzle_team_waves", "puzzle_hash_log2", "plain_grid_heaps"):
            assert got[k] == t[k]
        res = ctx.batch_solve(queries)
        for (kind, e), r in zip(expect, res):
            if kind == "grid":
                assert bool(r["found"]) == e["found"]
                if e["found"]:
                    assert (r["orth"], r["diag"]) == (e["orth"], e["diag"])
            else:
                assert r["found"] and r["steps"] == e
        for bad in (dict(got, puzzle_hash_log2=99), dict(got, grid_team_waves=24)):
            with pytest.raises(U.UAError):
                ctx.set_tuning(bad)
        ctx.set_tuning(dict(got, grid_team_waves=32))
        assert ctx.get_tuning()["grid_team_waves"] == 32
        for g in grids:
            g.free()
