"""Builds libmatflat_hip.so in-tree for gfx950 with hipcc.

Usage: ``python -m matflat_amd.build`` (also called by __graft_entry__.build()).

Flags:
* ``--offload-arch=gfx950``: MI355X only; there is no other target.
* ``-ffp-contract=off``: the exact kernels must never fuse a multiply and an
  add (MatFlat's results are defined without FMA).  Kernels that want FMA
  (the blocked mixed-precision path) re-enable it locally with
  ``#pragma clang fp contract(fast)``.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU
  instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmatflat_hip.so")
SOURCES = ["mfh_api.cpp", "mfh_exact.hip", "mfh_fast.hip", "mfh_gemm.hip", "mfh_adjoint.hip", "mfh_mixed.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *(extra or []), *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
