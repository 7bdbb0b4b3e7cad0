"""CPU checks of the C ABI boundary: libuastar.so loads (all device symbols resolved),
exports every entry point include/uastar.h declares, and the pure host helpers agree
with the oracle's input spec. No compute call needs a GPU here."""
import ctypes
import os
import re
import subprocess

import numpy as np

import uastar_amd as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "uastar.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ua_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_with_every_symbol_resolved():
    # RTLD_NOW: an undefined internal symbol fails here, not on the GPU box
    ctypes.CDLL(U.LIB_PATH, mode=os.RTLD_NOW)
    U.lib()


def test_exports_every_declared_entry_point():
    names = declared_entry_points()
    assert len(names) >= 20
    out = subprocess.check_output(["nm", "-D", "--defined-only", U.LIB_PATH], text=True)
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = U.lib()
    for n in names:
        assert getattr(L, n) is not None


def test_no_device_means_error_not_fallback():
    if U.device_count() > 0:
        return
    h = ctypes.c_void_p()
    assert U.lib().ua_create(0, None, ctypes.byref(h)) == -6  # UA_ERR_NO_DEVICE


def test_bitmap_pack_roundtrip(oracle):
    rng = np.random.default_rng(1)
    for H, W in [(1, 1), (3, 64), (5, 65), (17, 130), (9, 1000)]:
        g = np.where(rng.random(H * W) < 0.3, 0, 0xFF).astype(np.uint8)
        bm = U.bytes_to_bitmap(g, H, W)
        assert bm.shape == (H, (W + 63) // 64)
        assert np.array_equal(U.bitmap_to_bytes(bm, H, W), g)


def test_generator_spec_matches_oracle(oracle):
    H, W, seed, p = 40, 70, 99, 0.35
    g = oracle.grid_bernoulli(H, W, seed, p, 0, 0, H - 1, W - 1)
    for x in range(H):
        for y in range(W):
            if (x, y) in ((0, 0), (H - 1, W - 1)):
                continue
            assert U.cell_open(seed, p, W, x, y) == bool(g[x * W + y])


def test_key_to_pair_roundtrip():
    S = 97184015999
    for a, b in [(0, 0), (1, 0), (0, 1), (242, 879), (180000, 0), (3, 127000)]:
        assert U.key_to_pair(a * (1 << 36) + b * S) == (a, b)
