"""Autotuner of the batched kernels' launch shapes, with a persisted table.

It measures, on synthetic batches and with the library's per-launch HIP-event
timing (``Handle.set_timing``), every choice the library exposes per size
bucket and keeps the fastest:

* FAST Cholesky / SPD solve (``k_chol_fast``): wavefronts per matrix (1, 2 or
  4) for each 8-row bucket (``mfh_set_fast_waves``);
* exact Cholesky / SPD solve (``k_chol_packed``): the workgroup shape --
  lanes per matrix and rows per lane, i.e. how many matrices share a
  wavefront -- for each 16-row bucket (``mfh_set_chol_shape``), on orders
  uniform over the bucket and on the bucket's smallest order;
* the number of streams the FAST size buckets are forked over
  (``mfh_set_bucket_streams``), on the 100,000-matrix SPD-solve batch with
  n ~ U{10..100}; a bucket whose two best wavefront counts are within 5 %
  when timed alone is then decided inside that whole step, where the buckets
  run concurrently.

The table is written to ``matflat_amd/tuning/<arch>.json`` and applied to
every ``hip.Handle`` created afterwards on a device of that architecture
(``apply``).  The library's built-in defaults (``mfh_fast_default_waves``,
``chol_shape`` in mfh_exact.hip) hold the same choices for MI355X, so a C
caller that never loads the table runs the tuned shapes too.

    python -m matflat_amd.autotune [--count 8192] [--steps 5] [--no-write]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")

# lanes per matrix, rows per lane: the shapes k_chol_packed is built for
CHOL_SHAPES = [(8, 2), (16, 2), (16, 3), (32, 2), (64, 2), (64, 3)]
NBUCKETS, BUCKET_STEP = 9, 16


def table_path(arch: str) -> str:
    return os.path.join(TUNING_DIR, f"{arch}.json")


def chol_shapes_for(bucket: int):
    """Built shapes covering the bucket's rows whose wavefront fits the LDS
    (the library's own check, mfh_chol_shape_ok)."""
    rows = BUCKET_STEP * (bucket + 1)
    per_mat = rows * (rows - 1) // 2 + 6
    return [(g, s) for g, s in CHOL_SHAPES if g * s >= rows and per_mat * 8 * (64 // g) <= 160 * 1024]


def apply(handle, table: dict) -> None:
    """Set a handle's launch shapes from a tuning table."""
    for rb, w in table.get("fast_waves", {}).items():
        handle.set_fast_waves(int(rb), int(w))
    for b, (g, s) in table.get("chol_shape", {}).items():
        handle.set_chol_shape(int(b), int(g), int(s))
    if "bucket_streams" in table:
        handle.set_bucket_streams(int(table["bucket_streams"]))


def load(arch: str) -> dict | None:
    path = table_path(arch)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# ------------------------------------------------------------ measurement
def spd_batch(count: int, nmin: int, nmax: int, seed: int, pad8: bool = False):
    """count SPD matrices A = B B^T + n I (n ~ U{nmin..nmax}) in one device
    buffer (lda = n, or 8 ceil(n/8)); returns (n, a_ptrs, lda, x_ptrs, info,
    keep-alive buffers, pristine copy)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    ns = torch.randint(nmin, nmax + 1, (count,), generator=g, device="cuda").to(torch.int32)
    ldas = ((ns + 7) // 8 * 8) if pad8 else ns.clone()
    sizes = ldas.to(torch.int64) * ns.to(torch.int64)
    offs = torch.cumsum(sizes, 0) - sizes
    buf = torch.zeros(int(sizes.sum().item()), dtype=torch.float64, device="cuda")
    ns_host, ld_host, offs_host = ns.cpu().tolist(), ldas.cpu().tolist(), offs.cpu().tolist()
    by_n = {}
    for i, n in enumerate(ns_host):
        by_n.setdefault(n, []).append(i)
    for n, idx in by_n.items():
        bm = torch.randn((len(idx), n, n), dtype=torch.float64, device="cuda", generator=g)
        a = torch.bmm(bm, bm.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")
        ld = ld_host[idx[0]]
        full = torch.zeros((len(idx), n, ld), dtype=torch.float64, device="cuda")
        full[:, :, :n] = a  # row j of full = column j of A (A is symmetric)
        flat = full.reshape(len(idx), -1)
        for t, i in enumerate(idx):
            buf[offs_host[i]:offs_host[i] + n * ld] = flat[t]
    xoffs = torch.cumsum(ns.to(torch.int64), 0) - ns.to(torch.int64)
    x = torch.randn(int(ns.sum().item()), dtype=torch.float64, device="cuda", generator=g)
    a_ptrs = (buf.data_ptr() + 8 * offs).contiguous()
    x_ptrs = (x.data_ptr() + 8 * xoffs).contiguous()
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    return ns, a_ptrs, ldas, x_ptrs, info, (buf, x), (buf.clone(), x.clone())


def _timed(h, batch, nmax, steps, kind, rows):
    """Mean milliseconds of the launches of `kind` with row bound `rows`
    over `steps` SPD solves of the batch (restored before each call)."""
    import torch
    from . import _lib, hip
    ns, a_ptrs, lda, x_ptrs, info, (buf, x), (buf0, x0) = batch
    buf.copy_(buf0)
    x.copy_(x0)
    hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, nmax, handle=h)  # warm-up
    torch.cuda.synchronize()
    h.set_timing(True)
    for _ in range(steps):
        buf.copy_(buf0)
        x.copy_(x0)
        hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, nmax, handle=h)
    recs = h.timing()
    h.set_timing(False)
    if int((info != 0).sum().item()):
        raise RuntimeError(f"factorization failures while tuning (rows {rows})")
    ts = [ms for k, r, ms in recs if k == kind and r == rows]
    return sum(ts) / max(len(ts), 1)


def tune_fast_waves(h, count: int, steps: int, max_rb: int, log) -> tuple[dict, dict]:
    from . import _lib, hip
    h.set_math_mode(hip.MATH_FAST)
    best, times = {}, {}
    for rb in range(1, max_rb + 1):
        # SPD solve: n + 1 rows in (8(rb-1), 8rb]
        nmin, nmax = max(1, 8 * (rb - 1)), min(8 * rb - 1, _lib.SMALL_MAX_N)
        if nmin > nmax:
            continue
        batch = spd_batch(count, nmin, nmax, seed=rb)
        res = {}
        for w in (1, 2, 4):
            if (w == 2 and rb < 2) or (w == 4 and rb < 4):
                continue
            h.set_fast_waves(rb, w)
            res[w] = _timed(h, batch, nmax, steps, _lib.KERNEL_CHOL_FAST, 8 * rb)
        h.set_fast_waves(rb, 0)
        best[rb] = min(res, key=res.get)
        times[rb] = {str(w): round(t, 5) for w, t in res.items()}
        log(f"fast rb={rb:2d} n={nmin:3d}..{nmax:3d} " + " ".join(f"w{w}={t:.4f}ms" for w, t in res.items())
            + f" -> {best[rb]}")
        del batch
    return best, times


def tune_chol_shapes(h, count: int, steps: int, log) -> tuple[dict, dict]:
    """Each shape is timed on two batches of the bucket: orders uniform over
    it, and every matrix at its smallest order (where a shape with more
    lanes per matrix wastes the most).  The shape with the smallest sum of
    times relative to the best on each batch wins, so neither case is traded
    away for a small gain on the other."""
    from . import _lib, hip
    h.set_math_mode(hip.MATH_EXACT)
    best, times = {}, {}
    for b in range(NBUCKETS):
        shapes = chol_shapes_for(b)
        # SPD solve: n + 1 rows in (16b, 16(b+1)]
        nmin, nmax = max(1, 16 * b), min(16 * (b + 1) - 1, _lib.SMALL_MAX_N)
        if len(shapes) < 2 or nmin > nmax:
            continue
        rows = min(16 * (b + 1), nmax + 1)
        res = {}
        for tag, lo, hi, seed in (("uniform", nmin, nmax, 100 + b), ("smallest", nmin, nmin, 200 + b)):
            batch = spd_batch(count, lo, hi, seed=seed)
            for g, s in shapes:
                h.set_chol_shape(b, g, s)
                # the launch's row bound is that of the call's largest matrix
                res[(tag, g, s)] = _timed(h, batch, hi, steps, _lib.KERNEL_CHOL_EXACT, min(16 * (b + 1), hi + 1))
            del batch
        h.set_chol_shape(b, 0, 0)
        score = {}
        for g, s in shapes:
            score[(g, s)] = sum(res[(tag, g, s)] / min(res[(tag, gg, ss)] for gg, ss in shapes)
                                for tag in ("uniform", "smallest"))
        best[b] = list(min(score, key=score.get))
        times[b] = {f"{tag}:{g}x{s}": round(t, 5) for (tag, g, s), t in res.items()}
        log(f"exact bucket {b} rows {rows} " + " ".join(f"{tag[0]}:{g}x{s}={t:.4f}" for (tag, g, s), t in res.items())
            + f" -> {best[b]}")
    return best, times


def _step_ms(h, batch, steps: int) -> float:
    """Median milliseconds of one SPD-solve call over the whole batch (every
    bucket, forked over the handle's streams), inputs restored before each."""
    import torch
    from . import hip
    ns, a_ptrs, lda, x_ptrs, info, (buf, x), (buf0, x0) = batch
    hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, 100, handle=h)
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        buf.copy_(buf0)
        x.copy_(x0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        hip.dposv_vbatched(ns, a_ptrs, lda, x_ptrs, info, 100, handle=h)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return sorted(ts)[len(ts) // 2]


def refine_in_step(h, waves: dict, times: dict, batch, steps: int, log, margin: float = 0.05) -> dict:
    """Buckets whose two fastest choices are within `margin` alone are
    decided inside the whole SPD-solve step, where the buckets run
    concurrently on the forked streams."""
    out = dict(waves)
    for rb, tt in times.items():
        ranked = sorted(tt.items(), key=lambda kv: kv[1])
        if len(ranked) < 2 or ranked[1][1] > ranked[0][1] * (1 + margin) or 8 * rb > 104:
            continue
        res = {}
        for w, _ in ranked[:2]:
            h.set_fast_waves(rb, int(w))
            res[int(w)] = _step_ms(h, batch, steps)
        out[rb] = min(res, key=res.get)
        h.set_fast_waves(rb, out[rb])
        log(f"fast rb={rb} in the whole step: " + " ".join(f"w{w}={t:.4f}ms" for w, t in res.items())
            + f" -> {out[rb]}")
    return out


def tune_bucket_streams(h, batch, steps: int, log) -> tuple[int, dict]:
    from . import hip
    h.set_math_mode(hip.MATH_FAST)
    res = {}
    for k in (1, 2, 3, 4):
        h.set_bucket_streams(k)
        res[k] = _step_ms(h, batch, steps)
        log(f"bucket streams {k}: {res[k]:.4f} ms per step")
    best = min(res, key=res.get)
    h.set_bucket_streams(best)
    return best, {str(k): round(t, 5) for k, t in res.items()}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=8192, help="matrices per bucket batch")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--max-rb", type=int, default=16)
    ap.add_argument("--stream-count", type=int, default=100_000)
    ap.add_argument("--no-write", action="store_true")
    args = ap.parse_args(argv)

    import torch
    from . import hip

    def log(msg):
        print(f"[autotune] {msg}", file=sys.stderr, flush=True)

    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    h = hip.Handle(tuned=False)
    h.bind_stream()
    t0 = time.time()
    waves, wt = tune_fast_waves(h, args.count, args.steps, args.max_rb, log)
    shapes, st = tune_chol_shapes(h, args.count, args.steps, log)
    for rb, w in waves.items():
        h.set_fast_waves(rb, w)
    h.set_math_mode(hip.MATH_FAST)
    step_batch = spd_batch(args.stream_count, 10, 100, seed=7, pad8=True)
    streams, kt = tune_bucket_streams(h, step_batch, args.steps, log)
    waves = refine_in_step(h, waves, {rb: wt[rb] for rb in wt}, step_batch, args.steps, log)
    table = {
        "arch": arch,
        "fast_waves": {str(k): v for k, v in waves.items()},
        "chol_shape": {str(k): v for k, v in shapes.items()},
        "bucket_streams": streams,
        "measured_ms": {"fast_waves": {str(k): v for k, v in wt.items()},
                        "chol_shape": {str(k): v for k, v in st.items()},
                        "bucket_streams": kt},
        "workloads": {"per_bucket": f"{args.count} SPD solves per bucket, n uniform over the bucket, mean of "
                                    f"{args.steps} timed calls per choice (HIP events around each launch)",
                      "bucket_streams": f"{args.stream_count} SPD solves, n ~ U{{10..100}}, lda = 8 ceil(n/8), "
                                        f"median of {args.steps} calls"},
        "tuning_seconds": round(time.time() - t0, 1),
    }
    print(json.dumps(table))
    if not args.no_write:
        os.makedirs(TUNING_DIR, exist_ok=True)
        out = os.environ.get("MFH_TUNING_OUT") or table_path(arch)
        with open(out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        log(f"wrote {out}")
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
