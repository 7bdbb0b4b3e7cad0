"""Shared checks for the parity tests: path / move validators and input builders."""
from __future__ import annotations

import numpy as np

# utils.hpp:120-121 of uAstar
DX = [1, -1, 0, 0, 1, 1, -1, -1]
DY = [0, 0, 1, -1, 1, -1, 1, -1]


def grid_bytes(open_mask: np.ndarray) -> np.ndarray:
    """bool/int H x W (nonzero = open) -> the reference's byte-per-cell graph (0xFF / 0)."""
    return np.where(np.asarray(open_mask) != 0, 0xFF, 0).astype(np.uint8).reshape(-1)


def check_grid_path(path, graph: np.ndarray, H, W, sx, sy, ex, ey, orth, diag):
    """A path is valid when it starts and ends right, every step is one of the 8 moves out
    of an open cell (pathway/CPU-solver.cpp:64-70), and its step counts equal the cost."""
    path = [int(v) for v in path]
    assert path[0] == sx * W + sy, "path does not start at the start"
    assert path[-1] == ex * W + ey, "path does not end at the goal"
    g = np.asarray(graph).reshape(-1)
    n_orth = n_diag = 0
    for u, v in zip(path, path[1:]):
        ux, uy, vx, vy = u // W, u % W, v // W, v % W
        dx, dy = vx - ux, vy - uy
        assert max(abs(dx), abs(dy)) == 1, f"non-adjacent step {u}->{v}"
        d = [i for i in range(8) if DX[i] == dx and DY[i] == dy][0]
        assert g[u] & (1 << d), f"step out of a blocked cell {u}"
        assert 0 <= vx < H and 0 <= vy < W
        if abs(dx) + abs(dy) == 2:
            n_diag += 1
        else:
            n_orth += 1
    assert (n_orth, n_diag) == (orth, diag), f"path counts {(n_orth, n_diag)} != cost {(orth, diag)}"


def replay_moves(n, tiles, moves):
    """Applies blank moves (0..3 = (+1,0),(-1,0),(0,+1),(0,-1)) and returns the final board."""
    b = list(int(t) for t in np.asarray(tiles).reshape(-1))
    z = b.index(0)
    for m in moves:
        x, y = z // n + DX[m], z % n + DY[m]
        assert 0 <= x < n and 0 <= y < n, "move leaves the board"
        nz = x * n + y
        b[z], b[nz] = b[nz], b[z]
        z = nz
    return b


def goal_board(n):
    return list(range(1, n * n)) + [0]


def random_grid(rng, H, W, p):
    m = rng.random((H, W)) >= p
    return m


def check_grid_path_bitmap(path, bitmap: np.ndarray, H, W, sx, sy, ex, ey, orth, diag):
    """check_grid_path for large grids, vectorised over a row-padded bitmap (1 bit per
    cell, bit (y & 63) of word x*wpr + (y >> 6))."""
    p = np.asarray(path, np.int64)
    assert p[0] == sx * W + sy and p[-1] == ex * W + ey, "path endpoints"
    x, y = p // W, p % W
    dx, dy = np.diff(x), np.diff(y)
    step = np.maximum(np.abs(dx), np.abs(dy))
    assert np.all(step == 1), "non-adjacent step"
    assert np.all((x >= 0) & (x < H) & (y >= 0) & (y < W))
    bm = np.asarray(bitmap, np.uint64).reshape(H, -1)
    xs, ys = x[:-1], y[:-1]
    bits = (bm[xs, ys >> 6] >> (ys & 63).astype(np.uint64)) & np.uint64(1)
    assert np.all(bits == 1), "step out of a blocked cell"
    nd = int(np.sum(np.abs(dx) + np.abs(dy) == 2))
    assert (len(p) - 1 - nd, nd) == (orth, diag), "path counts differ from the cost"


def c4_grid_specs(n=4096, seed=4):
    """The grids of the batched configuration: side floor(exp(U(ln 256, ln 4096))),
    obstacles Bernoulli(p) with p ~ U(0.1, 0.4) from a per-grid seed, start and goal
    drawn uniformly among open cells (by the generator's cell rule)."""
    import math

    import uastar_amd as U
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(math.floor(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        p = float(rng.uniform(0.1, 0.4))
        gseed = int(rng.integers(1, 1 << 62))
        pts = []
        while len(pts) < 2:
            x, y = int(rng.integers(side)), int(rng.integers(side))
            if U.cell_open(gseed, p, side, x, y):
                pts.append((x, y))
        out.append(dict(H=side, W=side, seed=gseed, p=p, s=pts[0], g=pts[1]))
    return out
