"""Build libsep_amd.so (gfx950 only) in-tree with hipcc.

    python symmetriceigenproblem.jl_amd/build.py          # compile + link
    python symmetriceigenproblem.jl_amd/build.py --asm    # also dump gfx950 ISA to _build/

Every HIP source is compiled for gfx950 with packed-FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the library is linked as a plain
shared object whose only interface is the C ABI of ``include/sep.h``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsep_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = ["capi.hip", "apply_dpp.hip", "chase.hip", "batched.hip", "backward.hip",
           "tridiag.hip", "sort.hip"]

CFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src: str, asm: bool) -> str:
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)
    if asm:
        s = os.path.join(BUILD, src + ".gfx950.s")
        subprocess.run([HIPCC, *CFLAGS, "--cuda-device-only", "-S", os.path.join(CSRC, src),
                        "-o", s], check=True)
    return obj


def build(asm: bool = False, jobs: int | None = None) -> str:
    """Compile all sources (always) and link libsep_amd.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2) // 2))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, asm), SOURCES))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
                    "-L" + os.path.join(ROCM, "lib"), "-lrocblas"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--asm", action="store_true")
    a = ap.parse_args()
    print(build(asm=a.asm))
    sys.exit(0)
