"""ua_grid_solve_noalloc, the allocation-free plain grid solve (its buffers are one static
device array per GPU), against the allocating entry points ua_grid_solve + ua_grid_path on
the same queries: two calls in a row, the second on a smaller grid, each with the same
exact cost, a valid path of that cost, and a no-path case; a grid over the array's limit is
refused before any device work."""
import numpy as np
import pytest

import uastar_amd as U
from tests.helpers import check_grid_path_bitmap

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _check_pair(ctx, grid, H, W, s, g):
    ref = ctx.grid_solve(grid, *s, *g)
    ref_path = ctx.grid_path() if ref["found"] else None
    r, path = ctx.grid_solve_noalloc(grid, *s, *g)
    _say(f"noalloc {H}x{W}: {r['status']} ({r['orth']}, {r['diag']}) against ({ref['orth']}, {ref['diag']})")
    assert r["status"] == ref["status"] and bool(r["found"]) == bool(ref["found"])
    if ref["found"]:
        assert (r["orth"], r["diag"]) == (ref["orth"], ref["diag"])
        assert len(path) == len(ref_path) == r["orth"] + r["diag"] + 1
        check_grid_path_bitmap(path, grid.download(), H, W, *s, *g, r["orth"], r["diag"])
    else:
        assert path is None


def test_noalloc_matches_allocating_solve_twice_then_smaller():
    with U.Context(0) as ctx:
        cases = [(1000, 1000, 1, 0.30, (0, 0), (999, 999)),   # the 1000^2 configuration
                 (300, 700, 9, 0.25, (5, 3), (290, 650)),      # a second, smaller call
                 (64, 64, 4, 0.45, (0, 0), (63, 63))]          # dense obstacles
        for H, W, seed, p, s, g in cases:
            grid = ctx.grid_generate(H, W, seed, p, *s, *g)
            try:
                _check_pair(ctx, grid, H, W, s, g)
            finally:
                grid.free()
        # no path: the start boxed in
        m = np.ones((40, 40), bool)
        m[1, :3] = False
        m[:2, 2] = False
        grid = ctx.grid_from_bytes(np.where(m, 0xFF, 0).astype(np.uint8).reshape(-1), 40, 40)
        try:
            _check_pair(ctx, grid, 40, 40, (0, 0), (39, 39))
        finally:
            grid.free()
        # over the limit: refused with a clear error, nothing launched
        big = ctx.grid_generate(1025, 1024, 3, 0.3, 0, 0, 1024, 1023)
        try:
            with pytest.raises(U.UAError) as e:
                ctx.grid_solve_noalloc(big, 0, 0, 1024, 1023)
            assert e.value.status == -1 and "at most 1048576 cells" in str(e.value)
        finally:
            big.free()
