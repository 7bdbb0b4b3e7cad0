"""Builds uastar_amd/libuastar.so (HIP, gfx950) and the uastar CLI in-tree with hipcc.

    python -m uastar_amd.build            # library + CLI
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libuastar.so")
CLI = os.path.join(HERE, "uastar")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UASTAR_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code (the search is integer
# work; this keeps hipcc from forming v_pk_*_f32 anywhere).
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]

LIB_SOURCES = ["kernels.hip", "runtime.cpp", "step.cpp", "serve.cpp", "tune.cpp"]
CLI_SOURCES = ["cli/main.cpp", "cli/pathway.cpp", "cli/puzzle.cpp", "cli/cpu_solvers.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(verbose: bool = True) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    for src in LIB_SOURCES:
        obj = os.path.join(HERE, "_obj", os.path.basename(src) + ".o")
        _run([HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str | None:
    srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "cli"),
          *srcs, "-o", tmp, "-L", HERE, "-luastar", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build() -> None:
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)
