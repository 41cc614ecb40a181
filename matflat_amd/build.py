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

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmatflat_hip.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["mfh_api.cpp", "mfh_exact.hip", "mfh_fast.hip", "mfh_fast_posv.hip", "mfh_gemm.hip", "mfh_gemv.hip", "mfh_adjoint.hip", "mfh_mixed.hip", "mfh_general.hip", "mfh_gevd.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


# Per-source flags.  The FAST Cholesky kernels select a register-tile block
# by a uniform switch; SimplifyCFG's common-code sinking would merge the
# cases into one access at a computed index and push the tile to scratch.
SOURCE_FLAGS = {
    "mfh_fast.hip": ["-mllvm", "-simplifycfg-sink-common=false"],
    "mfh_fast_posv.hip": ["-mllvm", "-simplifycfg-sink-common=false"],
}


def _compile(src: str, obj: str, extra: list[str], verbose: bool) -> None:
    cmd = [HIPCC, *FLAGS, *SOURCE_FLAGS.get(os.path.basename(src), []), "-c", *extra, src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compiles every source to an object in parallel (the FAST and blocked
    kernels are template-heavy and take minutes each), then links."""
    extra = list(extra or [])
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_compile, s, o, extra, verbose) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
