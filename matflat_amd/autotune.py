"""Autotuner of the FAST Cholesky / SPD-solve kernels.

For every size bucket (matrices of up to 8*rb rows: n for the
factorization, n + 1 for the SPD solve) it times the kernel with one and
with two wavefronts per matrix on a synthetic batch, using the library's
per-launch event timing, and prints the faster choice per bucket.  The
library's built-in table (mfh_fast_default_waves in mfh_fast.hip) holds the
result for MI355X; a handle can override it with Handle.set_fast_waves.

    python -m matflat_amd.autotune [--count 8192] [--steps 5]
"""
from __future__ import annotations

import argparse
import json
import sys

import torch

from . import _lib, hip


def spd_batch(count: int, nmin: int, nmax: int, seed: int):
    """count SPD matrices A = B B^T + n I (n ~ U{nmin..nmax}), lda = n, in one
    device buffer; returns (n, a_ptrs, lda, x_ptrs, info, buffers)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    ns = torch.randint(nmin, nmax + 1, (count,), generator=g, device="cuda").to(torch.int32)
    sizes = (ns.to(torch.int64) ** 2)
    offs = torch.cumsum(sizes, 0) - sizes
    buf = torch.empty(int(sizes.sum().item()), dtype=torch.float64, device="cuda")
    ns_host = ns.cpu().tolist()
    offs_host = offs.cpu().tolist()
    for n in sorted(set(ns_host)):
        idx = [i for i, v in enumerate(ns_host) if v == n]
        bm = torch.randn((len(idx), n, n), dtype=torch.float64, device="cuda", generator=g)
        a = torch.bmm(bm, bm.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")
        for t, i in enumerate(idx):
            buf[offs_host[i]:offs_host[i] + n * n] = a[t].reshape(-1)
    xoffs = torch.cumsum(ns.to(torch.int64), 0) - ns.to(torch.int64)
    x = torch.randn(int(ns.sum().item()), dtype=torch.float64, device="cuda", generator=g)
    a_ptrs = (buf.data_ptr() + 8 * offs).contiguous()
    x_ptrs = (x.data_ptr() + 8 * xoffs).contiguous()
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    return ns, a_ptrs, ns.clone(), x_ptrs, info, (buf, x)


def time_bucket(h: hip.Handle, rb: int, waves: int, batch, nmax: int, steps: int) -> float:
    ns, a_ptrs, lda, x_ptrs, info, _ = batch
    h.set_fast_waves(rb, waves)
    hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, nmax, handle=h)  # warm-up
    torch.cuda.synchronize()
    h.set_timing(True)
    for _ in range(steps):
        hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, nmax, handle=h)
    recs = h.timing()
    h.set_timing(False)
    if int((info != 0).sum().item()):
        raise RuntimeError(f"factorization failures in bucket rb={rb}")
    ts = [ms for kind, rows, ms in recs if kind == _lib.KERNEL_CHOL_FAST and rows == 8 * rb]
    return sum(ts) / len(ts)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--max-rb", type=int, default=16)
    args = ap.parse_args(argv)
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST)
    table = {}
    for rb in range(1, args.max_rb + 1):
        # SPD solve: n + 1 rows in (8(rb-1), 8rb]; n <= MFH_SMALL_MAX_N
        nmin, nmax = max(1, 8 * (rb - 1)), min(8 * rb - 1, _lib.SMALL_MAX_N)
        batch = spd_batch(args.count, nmin, nmax, seed=rb)
        res = {}
        for waves in [w for w in (1, 2, 4) if w == 1 or (w == 2 and rb >= 2) or (w == 4 and rb >= 4)]:
            res[waves] = time_bucket(h, rb, waves, batch, nmax, args.steps)
        h.set_fast_waves(rb, 0)
        best = min(res, key=res.get)
        table[rb] = {"best": best, **{f"ms_w{w}": round(t, 5) for w, t in res.items()}}
        print(f"rb={rb:2d} n={nmin:3d}..{nmax:3d} " + " ".join(f"w{w}={t:.4f}ms" for w, t in res.items())
              + f" -> {best}", file=sys.stderr, flush=True)
        del batch
    print(json.dumps(table))
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
