"""Benchmark: variable-size batched FP64 SPD solve (Cholesky + two triangular
solves) on one MI355X.

Workload (one step): 100,000 SPD matrices, order n ~ U{10..100}, leading
dimension 8*ceil(n/8), A = B B^T + n I with B ~ N(0,1) (upper triangle
stored, strict lower triangle NaN: it is never read), one N(0,1) right-hand
side each.  A step is one replay of the hipGraph that holds the whole batch
(device bucketing + the solve kernels) on a pristine copy of the inputs,
resident in HBM.  The solve works in place, so the timed steps rotate over
working copies that are restored from the pristine batch between cycles,
outside the timing events (two copies of the 3.1 GB batch: no input is in
the 256 MB last-level cache when its step starts); the outputs of the last
timed step are checked (every info zero, residuals of a sample).

Prints one JSON line (see README for the fields).  --gpus N with N > 1 exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "solves/s, batched var-size LU/Chol n≤512: FP32+IR vs FP64 at equal bkwd err"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _plain(v):
    """JSON form of the numpy scalars the result dictionaries may hold."""
    if isinstance(v, np.generic):
        return v.item()
    raise TypeError(f"{type(v).__name__} is not JSON serializable")


def host_cpu() -> dict:
    """The host CPU the baselines run on and the threads they use: the OpenMP
    thread count the environment sets (the host's CPU share for this GPU),
    capped by the cores this process may run on; every allowed core when the
    environment sets none."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "").strip()
    threads = min(allowed, int(env)) if env.isdigit() and int(env) > 0 else allowed
    return {"cpu_model": model, "logical_cores_visible": os.cpu_count(), "cores_allowed": allowed,
            "threads": threads, "omp_num_threads": env or None}


def cpu_entry(value: float, unit: str, sample: str, reps: int) -> dict:
    h = host_cpu()
    return {"value": round(value, 1), "unit": unit, "cores": h["threads"], "kind": "port", "sample": sample,
            "runs": reps, "cpu_model": h["cpu_model"], "cores_allowed": h["cores_allowed"],
            "logical_cores_visible": h["logical_cores_visible"],
            "thread_policy": "OMP_NUM_THREADS (the host's CPU share for this GPU) capped by the affinity mask"}


def timed_cpu(run, max_seconds: float, min_reps: int = 3, max_reps: int = 5):
    """Median wall time of run() (each call does its own untimed setup and
    returns its timed seconds) over min_reps..max_reps runs, stopping once
    max_seconds have passed."""
    elapsed = []
    t_start = time.perf_counter()
    while len(elapsed) < max_reps:
        elapsed.append(run())
        if len(elapsed) >= min_reps and time.perf_counter() - t_start > max_seconds:
            break
    return float(np.median(elapsed)), len(elapsed)


def rotated_calls(call, restore, copies: int, steps: int, warmup: int, stream):
    """Per-call durations (ms) of `steps` timed calls, each bracketed by its
    own HIP events on `stream`.  The calls cycle over `copies` working copies
    of the inputs: a cycle first restores every copy (restore(i), untimed and
    outside the events; nothing when restore is None), then calls copy 0, 1,
    ... in the order they were restored, so the other copies' bytes pass
    through the last-level cache between a copy's restore and its call.
    Returns (per-call ms list, index of the copy the last timed call used)."""
    import torch

    def cycle(n_calls, timed):
        if restore is not None:
            for i in range(copies):
                restore(i)
        evs = []
        for i in range(n_calls):
            if timed:
                s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s0.record(stream)
                call(i)
                s1.record(stream)
                evs.append((s0, s1))
            else:
                call(i)
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in evs]

    done = 0
    while done < warmup:
        k = min(copies, warmup - done)
        cycle(k, False)
        done += k
    times = []
    last = 0
    while len(times) < steps:
        k = min(copies, steps - len(times))
        times += cycle(k, True)
        last = k - 1
    return times, last


def make_spd_batch(count: int, nmin: int, nmax: int, seed: int, device: str = "cuda"):
    """Generates the batch on the GPU.  Returns (buf, offsets, ns, ldas, b)."""
    import torch

    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count).astype(np.int32)
    ldas = ((ns + 7) // 8 * 8).astype(np.int32)
    # memory grouped by size, batch order shuffled
    order = np.argsort(ns, kind="stable")
    sizes = ldas.astype(np.int64) * ns
    offsets = np.zeros(count, dtype=np.int64)
    offsets[order] = np.concatenate([[0], np.cumsum(sizes[order])[:-1]])
    total = int(sizes.sum())
    buf = torch.full((total,), float("nan"), dtype=torch.float64, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    for n in range(nmin, nmax + 1):
        idx = np.flatnonzero(ns == n)
        if idx.size == 0:
            continue
        lda = int((n + 7) // 8 * 8)
        c = idx.size
        bm = torch.randn((c, n, n), dtype=torch.float64, device=device, generator=g)
        a = torch.bmm(bm, bm.transpose(1, 2).contiguous())
        a += n * torch.eye(n, dtype=torch.float64, device=device)
        # column-major slab [c][col][row(lda)], upper triangle only
        slab = torch.full((c, n, lda), float("nan"), dtype=torch.float64, device=device)
        upper = torch.triu(torch.ones(n, n, dtype=torch.bool, device=device))  # [row, col], row <= col
        at = a.transpose(1, 2)  # at[b, col, row] = a[b, row, col]
        slab[:, :, :n] = torch.where(upper.T.unsqueeze(0), at, torch.full_like(at, float("nan")))
        start = int(offsets[idx].min())
        buf[start:start + c * n * lda] = slab.reshape(-1)
        del bm, a, slab, at
    xoff = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]])
    b = torch.randn(int(ns.sum()), dtype=torch.float64, device=device, generator=g)
    return buf, offsets, ns, ldas, xoff, b


def alg_bytes_posv(ns: np.ndarray) -> float:
    n = ns.astype(np.float64)
    return float(np.sum(8.0 * (n * (n + 1) / 2 + n * n + 2 * n) + 4.0))


def kernel_pass(handle, mode, batch, nmax, ns, steps, stream):
    """Times every kernel launch of `steps` eager batched solves, each on a
    pristine working copy (restored between cycles, outside the kernels),
    with HIP events recorded by the library on the stream it launches on.
    Returns one entry per distinct launch: kind, row bound, matrices it
    processes, their algorithmic bytes and the mean duration."""
    from matflat_amd import _lib, hip

    handle.set_math_mode(mode)
    handle.set_timing(True)

    def call(i):
        hip.dposv_vbatched(batch.n_t, batch.a_ptrs[i], batch.lda_t, batch.x_ptrs[i], batch.info[i], nmax,
                           handle=handle)

    rotated_calls(call, batch.restore, batch.copies, steps, 0, stream)
    recs = handle.timing()
    handle.set_timing(False)
    per: dict = {}
    for kind, rows, t in recs:
        per.setdefault((kind, rows), []).append(t)
    keys = ns.astype(np.int64) + 1  # rows of an SPD solve: n + the right-hand side
    out = []
    for kind in sorted({k for k, _ in per}):
        bounds = sorted(r for k, r in per if k == kind)
        lo = 0
        for r in bounds:
            times = per[(kind, r)]
            entry = {"kernel": _lib.KERNEL_NAMES.get(kind, str(kind)), "kind": kind, "rows": r,
                     "launches": len(times), "avg_ms": float(np.mean(times))}
            if kind != _lib.KERNEL_SORT:
                sel = ns[(keys > lo) & (keys <= r)]
                entry["matrices"] = int(sel.size)
                entry["alg_bytes"] = alg_bytes_posv(sel)
                entry["gbs"] = entry["alg_bytes"] / (entry["avg_ms"] / 1e3) / 1e9
                lo = r
            out.append(entry)
    for e in out:
        log(f"  {e['kernel']:14s} rows<={e['rows']:3d} x{e['launches']} avg {e['avg_ms']:.4f} ms"
            + (f"  {e['matrices']} matrices {e['gbs']:.1f} GB/s" if "gbs" in e else ""))
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_chol_fast_posv_100k_spd_traffic.md")


def measured_traffic(rows: int, default_workload: bool):
    """HBM bytes per launch of the solve kernel with this row bound, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this benchmark's
    default workload (calibrated against a copy kernel); None when the
    workload differs or no figure is recorded."""
    if not default_workload or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        parts = dict(kv.split("=") for kv in line.split()[1:]) if line.startswith("traffic:") else {}
        if parts and int(parts["rows"]) == rows:
            return int(parts["bytes"])
    return None


def roofline_of(kern, ns, default_workload: bool):
    """Roofline entry for the dominant solve kernel (largest mean duration)."""
    solve = [e for e in kern if "gbs" in e]
    dom = max(solve, key=lambda e: e["avg_ms"])
    achieved = dom["gbs"]
    step_ms = sum(e["avg_ms"] * e["launches"] for e in kern) / max(1, max(e["launches"] for e in kern))
    all_bytes = alg_bytes_posv(ns)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom["rows"], default_workload),
            "traffic_source": "profiles/k_chol_fast_posv_100k_spd_traffic.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE "
                              "x1.0, scaled by a 256 MiB copy)",
            "kernel": f"{dom['kernel']} (row bound {dom['rows']}: {dom['matrices']} matrices)",
            "kernel_avg_ms": round(dom["avg_ms"], 5), "alg_bytes_per_launch": dom["alg_bytes"],
            "all_kernels": {"eager_step_ms": round(step_ms, 5), "alg_bytes_per_step": all_bytes,
                            "achieved_gbs": round(all_bytes / (step_ms / 1e3) / 1e9, 1),
                            "launches": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in e.items()}
                                         for e in kern]}}


def copy_calibration(mib: int = 256):
    """One device-to-device copy of a known size, for scaling HBM counters."""
    import torch

    src = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: {mib} MiB read + {mib} MiB written")


MIXED_SIZES = (128, 256, 384, 512)


def make_conditioned_batch(count: int, seed: int, chunk: int = 200):
    """The mixed-precision workload: n uniform over {128, 256, 384, 512},
    kappa = 10^U(2,6), sigma_i = kappa^(-i/(n-1)), A = U diag(sigma) V^T with
    Haar U, V (QR of N(0,1) with the R-diagonal sign fix), lda = n, b ~ N(0,1),
    all in FP64 on the GPU.  Returns (buf, offsets, ns, b, boffsets)."""
    import torch

    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(MIXED_SIZES), size=count).astype(np.int32)
    kappa = 10.0 ** rng.uniform(2.0, 6.0, size=count)
    order = np.argsort(ns, kind="stable")
    sizes = ns.astype(np.int64) ** 2
    offsets = np.zeros(count, dtype=np.int64)
    offsets[order] = np.concatenate([[0], np.cumsum(sizes[order])[:-1]])
    buf = torch.empty(int(sizes.sum()), dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)

    def haar(c, n):
        q, r = torch.linalg.qr(torch.randn((c, n, n), dtype=torch.float64, device="cuda", generator=g))
        return q * torch.sign(torch.diagonal(r, dim1=1, dim2=2)).unsqueeze(1)

    for n in MIXED_SIZES:
        idx = np.flatnonzero(ns == n)
        log(f"  generating {idx.size} matrices of order {n}")
        for s0 in range(0, idx.size, chunk):
            sub = idx[s0:s0 + chunk]
            c = sub.size
            expo = torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1)
            sig = torch.from_numpy(kappa[sub]).to("cuda").unsqueeze(1) ** (-expo.unsqueeze(0))
            u, v = haar(c, n), haar(c, n)
            a = torch.bmm(u * sig.unsqueeze(1), v.transpose(1, 2).contiguous())
            start = int(offsets[sub[0]])
            buf[start:start + c * n * n] = a.transpose(1, 2).contiguous().reshape(-1)  # column-major
            del u, v, a
            torch.cuda.synchronize()
            log(f"    order {n}: {s0 + c} of {idx.size}")
    boff = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]])
    b = torch.randn(int(ns.sum()), dtype=torch.float64, device="cuda", generator=g)
    return buf, offsets, ns, b, boff


MFMA_PROFILE = os.path.join(ROOT, "profiles", "k_gesv_8000_n128-512_mfma.md")


def mfma_busy(kernel: str):
    """MFMA busy fraction (SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x 1024
    SIMDs)) of a factorization kernel of the 8,000-matrix mixed-precision
    workload, from the committed rocprofv3 --pmc summary; None when absent."""
    if not os.path.exists(MFMA_PROFILE):
        return None
    import re

    for line in open(MFMA_PROFILE):
        m = re.match(r"mfma_busy: kernel=(.+) frac=(\S+)\s*$", line)
        if m and m.group(1) == kernel:
            return float(m.group(2))
    return None


def mixed_cpu_baseline(buf, offsets, ns, b, boff, per_size=100, max_seconds=12.0):
    """The oracle's LU (Factorization.Lu) + LU solve over a bounded sample of
    the mixed-precision batch: the first `per_size` matrices of each order,
    one threaded batch per order, OpenMP dynamic over the host's threads."""
    import oracle

    threads = host_cpu()["threads"]
    groups = []
    b_host = b.cpu().numpy()
    for n in MIXED_SIZES:
        idx = np.flatnonzero(ns == n)[:per_size]
        if idx.size == 0:
            continue
        o = int(offsets[idx[0]])
        assert np.array_equal(offsets[idx], o + n * n * np.arange(idx.size))  # contiguous per order
        a = buf[o:o + idx.size * n * n].cpu().numpy()
        rhs = np.concatenate([b_host[boff[j]:boff[j] + n] for j in idx])
        groups.append((n, idx.size, a, rhs))
    count = sum(g[1] for g in groups)

    def run():
        work = [(n, c, a.copy(), rhs) for n, c, a, rhs in groups]
        t0 = time.perf_counter()
        for n, c, a, rhs in work:
            oracle.dgesv_batch(c, n, n, n * n, a, rhs, threads=threads)
        return time.perf_counter() - t0

    med, reps = timed_cpu(run, max_seconds, min_reps=3, max_reps=5)
    return cpu_entry(count / med, "solves/s",
                     f"{count} of the 8,000 matrices (the first {per_size} of each order in {list(MIXED_SIZES)}, the "
                     "batch's own size mix), FP64 LU with partial pivoting + LU solve in MatFlat's order "
                     "(oracle/matflat_oracle.c, Factorization.Lu.cs:296-350 and Factorization.LuInverse.cs:169-192), "
                     f"OpenMP dynamic over {threads} threads, median of {reps} runs", reps)


def save_conditioned_batch(path: str, count: int = 8000, seed: int = 3) -> None:
    """Generates the mixed-precision workload and writes it to `path` (a
    directory) as .npy files, for runs that must not generate it themselves
    (a counter-collecting profiler run, which fails inside torch's QR)."""
    buf, offsets, ns, b, boff = make_conditioned_batch(count, seed)
    os.makedirs(path, exist_ok=True)
    np.save(os.path.join(path, "a.npy"), buf.cpu().numpy())
    np.save(os.path.join(path, "b.npy"), b.cpu().numpy())
    np.savez(os.path.join(path, "meta.npz"), offsets=offsets, ns=ns, boff=boff, seed=seed)
    log(f"saved the {count}-matrix batch to {path}")


def load_conditioned_batch(path: str):
    import torch

    meta = np.load(os.path.join(path, "meta.npz"))
    buf = torch.from_numpy(np.load(os.path.join(path, "a.npy"), mmap_mode="r")[:].copy()).to("cuda")
    b = torch.from_numpy(np.load(os.path.join(path, "b.npy"))).to("cuda")
    return buf, meta["offsets"], meta["ns"], b, meta["boff"]


def mixed_leg(steps: int, warmup: int, count: int = 8000, seed: int = 3, cpu: bool = True, batch_dir=None):
    """FP32/BF16-factor + FP64-refinement LU solve against the plain FP64
    blocked LU solve on the same batch, both to backward error <= 1e-15 n.
    The inputs are read, never written, so every step sees the same bytes;
    the 7.9 GB batch is 30x the last-level cache."""
    import torch

    from matflat_amd import _lib, hip

    if batch_dir:
        log(f"mixed precision: loading the batch from {batch_dir}")
        buf, offsets, ns, b, boff = load_conditioned_batch(batch_dir)
        count = ns.size
    else:
        log(f"mixed precision: generating {count} matrices n in {MIXED_SIZES}, cond 1e2..1e6")
        buf, offsets, ns, b, boff = make_conditioned_batch(count, seed)
    torch.cuda.synchronize()
    dev = torch.device("cuda")
    n_t = torch.from_numpy(ns).to(dev)
    a_ptrs = (buf.data_ptr() + 8 * torch.from_numpy(offsets).to(dev)).contiguous()
    b_ptrs = (b.data_ptr() + 8 * torch.from_numpy(boff).to(dev)).contiguous()
    x = torch.zeros_like(b)
    x_ptrs = (x.data_ptr() + 8 * torch.from_numpy(boff).to(dev)).contiguous()
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    iters = torch.zeros(count, dtype=torch.int32, device=dev)
    berr = torch.zeros(count, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    max_n = max(MIXED_SIZES)

    def call(mixed):
        if mixed:
            hip.dgesv_mixed_vbatched(n_t, a_ptrs, n_t, b_ptrs, x_ptrs, 10, 1e-15, iters, berr, info, max_n,
                                     handle=handle)
        else:
            hip.dgesv_vbatched(n_t, a_ptrs, n_t, b_ptrs, x_ptrs, berr, info, max_n, handle=handle)

    res = {}
    nf = ns.astype(np.float64)
    flops = float(np.sum(2.0 / 3.0 * nf ** 3 + 1.5 * nf ** 2))
    factor_flops = float(np.sum(2.0 / 3.0 * nf ** 3))
    for name, mixed in (("fp64", False), ("mixed", True)):
        times, _ = rotated_calls(lambda i: call(mixed), None, 1, steps, warmup, stream)
        ms = float(np.mean(times))
        inf = info.cpu().numpy()
        be = berr.cpu().numpy()
        ratio = be / ns
        entry = {"ms_per_step": round(ms, 4), "solves_per_s": round(count / (ms / 1e3), 1),
                 "hpl_tflops": round(flops / (ms / 1e3) / 1e12, 2),
                 "max_berr_over_n": float(ratio.max()), "all_info_zero": bool(np.all(inf == 0)),
                 "meets_1e-15n": bool(np.all(ratio <= 1e-15))}
        if mixed:
            it = iters.cpu().numpy()
            entry.update({"mean_iters": round(float(it.mean()), 3), "max_iters": int(it.max())})
        # one call with the library's per-launch events: the factorization kernel and the solve steps
        handle.set_timing(True)
        call(mixed)
        recs = handle.timing()
        handle.set_timing(False)
        fkind = _lib.KERNEL_GESV_MIXED if mixed else _lib.KERNEL_GESV_F64
        fact = [t for k, _, t in recs if k == fkind]
        step = [t for k, _, t in recs if k == _lib.KERNEL_GESV_STEP]
        kname = "k_gesv<float, true>" if mixed else "k_gesv<double, false>"
        fms = float(np.mean(fact))
        tf = factor_flops / (fms / 1e3) / 1e12
        busy = mfma_busy(kname)
        entry["kernels"] = {"factor_ms": round(fms, 4), "solve_step_launches": len(step),
                            "solve_steps_ms": round(float(np.sum(step)), 4)}
        entry["roofline"] = {"bound": "mfma", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                             "frac": round(tf / 78.6, 4), "traffic": None,
                             "kernel": f"{kname} (factorization of all {count} matrices, one launch)",
                             "kernel_avg_ms": round(fms, 5), "alg_flops_per_launch": factor_flops,
                             "peak_basis": "FP64 dense, 78.6 TFLOP/s; flops are the LU's 2/3 n^3 (HPL count) "
                                           "whatever precision computes them",
                             "mfma_busy": busy,
                             "mfma_busy_source": ("profiles/k_gesv_8000_n128-512_mfma.md (rocprofv3 --pmc "
                                                  "SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x 1024 SIMDs))"
                                                  if busy is not None else None)}
        log(f"  {name}: {ms:.3f} ms/step, {entry['solves_per_s']:.0f} solves/s, {entry['hpl_tflops']} TFLOP/s (HPL "
            f"count), max berr/n {entry['max_berr_over_n']:.2e}; factor kernel {fms:.3f} ms ({tf:.2f} TFLOP/s), "
            f"{len(step)} solve launches {float(np.sum(step)):.3f} ms" +
            (f", iters mean {entry['mean_iters']} max {entry['max_iters']}" if mixed else ""))
        res[name] = entry
    handle.close()
    res["speedup_mixed_over_fp64"] = round(res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"], 3)
    res["roofline"] = dict(res["mixed"]["roofline"])
    res["workload"] = (f"{count} LU solves, n uniform over {list(MIXED_SIZES)}, A = U diag(s) V^T (Haar U, V), "
                       "cond 10^U(2,6), b ~ N(0,1); refine to ||b-Ax||/(||A|| ||x|| + ||b||) <= 1e-15 n or 10 "
                       "iterations")
    res["seed"] = seed
    if cpu:
        log("  mixed-precision workload: cpu baseline")
        res["cpu_baseline"] = mixed_cpu_baseline(buf, offsets, ns, b, boff)
        log(f"  cpu: {res['cpu_baseline']['value']:.1f} solves/s")
    del buf, b, x
    torch.cuda.empty_cache()
    return res


def timed_calls(call, restore, steps: int, warmup: int, stream):
    """Mean ms of `call` over `steps` runs, each bracketed by its own HIP events
    on `stream`; `restore` (untimed, may be None) runs before every call."""
    import torch

    for _ in range(warmup):
        if restore:
            restore()
        call()
    torch.cuda.synchronize()
    total = 0.0
    for _ in range(steps):
        if restore:
            restore()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        call()
        stop.record(stream)
        stop.synchronize()
        total += start.elapsed_time(stop)
    return total / steps


def hbm_entry(alg_bytes: float, ms: float) -> dict:
    gbs = alg_bytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "alg_bytes": alg_bytes}


def lu_leg(steps: int, warmup: int, count: int = 1024, n: int = 64, seed: int = 1, copies: int = 16,
           cpu: bool = True):
    """Exact (bit-identical to MatFlat) partial-pivot LU of `count` column-major
    n x n matrices, lda = n, N(0,1) entries; then LuInverse and the LU solve
    of the same factors.  Each timed call works on its own pristine working
    copy: the calls rotate over `copies` copies (16 x 33.5 MB = 537 MB, twice
    the last-level cache), restored between cycles outside the events."""
    import torch

    from matflat_amd import hip

    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    per = count * n * n
    src = torch.randn(per, dtype=torch.float64, device=dev, generator=g)
    work = torch.empty((copies, per), dtype=torch.float64, device=dev)
    piv = torch.zeros((copies, count * n), dtype=torch.int32, device=dev)
    sign = torch.zeros((copies, count), dtype=torch.int32, device=dev)
    info = torch.zeros((copies, count), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    handle.set_math_mode(hip.MATH_EXACT)

    def restore_src(i):
        work[i].copy_(src)
        info[i].fill_(-99)

    times, last = rotated_calls(
        lambda i: hip.dgetrf_strided_batched(count, n, n, work[i], n, n * n, piv[i], n, sign[i], info[i],
                                             handle=handle), restore_src, copies, steps, warmup, stream)
    ms = float(np.mean(times))
    ok = bool(torch.all(info[last] == 0).item())
    got_lu, got_piv = work[last].cpu().numpy(), piv[last].cpu().numpy()
    # Factorization.LuInverse of the factored batch: every copy restored to the factors
    lu_src = work[last].clone()
    piv_src = piv[last].clone()
    idx = torch.arange(count, dtype=torch.int64, device=dev)
    a_ptrs = [(work[i].data_ptr() + 8 * n * n * idx).contiguous() for i in range(copies)]
    piv_ptrs = [(piv[i].data_ptr() + 4 * n * idx).contiguous() for i in range(copies)]
    n_t = torch.full((count,), n, dtype=torch.int32, device=dev)
    inv_info = torch.zeros((copies, count), dtype=torch.int32, device=dev)

    def restore_lu(i):
        work[i].copy_(lu_src)
        piv[i].copy_(piv_src)
        inv_info[i].fill_(-99)

    inv_times, inv_last = rotated_calls(
        lambda i: hip.dgetri_vbatched(n_t, a_ptrs[i], n_t, piv_ptrs[i], inv_info[i], n, handle=handle),
        restore_lu, copies, steps, warmup, stream)
    ms_inv = float(np.mean(inv_times))
    inv_ok = bool(torch.all(inv_info[inv_last] == 0).item())
    # LU solve (the LuInverse composition for one right-hand side): read-only factors, rotated
    for i in range(copies):
        restore_lu(i)
    bvec = torch.randn((copies, count * n), dtype=torch.float64, device=dev, generator=g)
    xvec = torch.zeros_like(bvec)
    b_ptrs = [(bvec[i].data_ptr() + 8 * n * idx).contiguous() for i in range(copies)]
    x_ptrs = [(xvec[i].data_ptr() + 8 * n * idx).contiguous() for i in range(copies)]
    rs_info = torch.zeros((copies, count), dtype=torch.int32, device=dev)
    rs_times, rs_last = rotated_calls(
        lambda i: hip.dgetrs_vbatched(n_t, a_ptrs[i], n_t, piv_ptrs[i], b_ptrs[i], x_ptrs[i], rs_info[i], n,
                                      handle=handle), None, copies, steps, warmup, stream)
    ms_rs = float(np.mean(rs_times))
    rs_ok = bool(torch.all(rs_info[rs_last] == 0).item())
    handle.close()
    alg = count * (16.0 * n * n + 4 * n + 4)
    res = {"workload": f"{count} LU factorizations (partial pivoting), n = {n}, lda = {n}, N(0,1), exact mode "
                       "(bit-identical to MatFlat's Factorization.Lu)",
           "working_copies": copies, "working_set_mb": round(copies * per * 8 / 1e6, 1),
           "ms_per_call": round(ms, 5), "ms_per_call_median": round(float(np.median(times)), 5),
           "factorizations_per_s": round(count / (ms / 1e3), 1),
           "all_info_zero": ok, "roofline": hbm_entry(alg, ms),
           "inverse": {"workload": f"Factorization.LuInverse of the same {count} factored matrices, exact mode "
                                   "(bit-identical)",
                       "ms_per_call": round(ms_inv, 5), "inverses_per_s": round(count / (ms_inv / 1e3), 1),
                       "all_info_zero": inv_ok, "roofline": hbm_entry(alg, ms_inv)},
           "solve": {"workload": f"LU solve (Factorization.LuInverse.cs:169-192 for one right-hand side) with the same "
                                 f"{count} factors, exact mode (bit-identical)",
                     "ms_per_call": round(ms_rs, 5), "solves_per_s": round(count / (ms_rs / 1e3), 1),
                     "all_info_zero": rs_ok,
                     "roofline": hbm_entry(count * (8.0 * n * n + 4 * n + 16 * n + 4), ms_rs)}}
    if cpu:
        import oracle

        threads = host_cpu()["threads"]
        a_host = src.cpu().numpy()
        out = {}

        def run():
            a = a_host.copy()
            t0 = time.perf_counter()
            out["piv"], _ = oracle.dgetrf_batch(count, n, n, n, n * n, a, threads=threads)
            dt = time.perf_counter() - t0
            out["lu"] = a
            return dt

        med, reps = timed_cpu(run, 3.0, min_reps=5, max_reps=50)
        res["cpu_baseline"] = cpu_entry(count / med, "factorizations/s",
                                        f"the full {count}-matrix batch, Factorization.Lu in MatFlat's order "
                                        f"(oracle/matflat_oracle.c), OpenMP dynamic over {threads} threads, "
                                        f"median of {reps} runs", reps)
        # the timed call's outputs against the oracle's: MatFlat's pivots and bits
        res["output_check"] = {"pivots_bit_exact": bool(np.array_equal(got_piv + 0, out["piv"])),
                               "factors_bit_exact": bool(np.array_equal(got_lu, out["lu"]))}
    log(f"  LU n={n} x{count}: {ms:.4f} ms/call, {res['factorizations_per_s']:.0f} LU/s, "
        f"{res['roofline']['achieved']} GB/s; inverse {ms_inv:.4f} ms/call; solve {ms_rs:.4f} ms/call"
        + (f"; cpu {res['cpu_baseline']['value']:.0f} LU/s, check {res['output_check']}" if cpu else ""))
    return res


def adjoint_leg(steps: int, warmup: int, count: int = 50_000, n: int = 32, seed: int = 4, copies: int = 2,
                cpu: bool = True):
    """Forward (exact-mode SPD solve) and hand-written backward of the batched
    SPD solve, n = 32, A = B B^T + n I.  The forward works in place, so its
    calls rotate over `copies` pristine working copies (2 x 410 MB) restored
    between cycles; the backward reads the forward's outputs of each copy."""
    import torch

    from matflat_amd import hip

    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    bm = torch.randn((count, n, n), dtype=torch.float64, device=dev, generator=g)
    src = torch.bmm(bm, bm.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device=dev)
    del bm
    b0 = torch.randn((count, n), dtype=torch.float64, device=dev, generator=g)
    w = torch.randn((count, n), dtype=torch.float64, device=dev, generator=g)
    a = [torch.empty_like(src) for _ in range(copies)]
    x = [torch.empty_like(b0) for _ in range(copies)]
    abar, bbar = torch.empty_like(src), torch.empty_like(b0)
    ns = torch.full((count,), n, dtype=torch.int32, device=dev)
    step = torch.arange(count, dtype=torch.int64, device=dev)

    def ptrs(t, per):
        return (t.data_ptr() + 8 * per * step).contiguous()

    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    handle.set_math_mode(hip.MATH_EXACT)
    pa = [ptrs(t, n * n) for t in a]
    px = [ptrs(t, n) for t in x]
    pw, pab, pbb = ptrs(w, n), ptrs(abar, n * n), ptrs(bbar, n)

    def restore(i):
        a[i].copy_(src)
        x[i].copy_(b0)

    ft, _ = rotated_calls(lambda i: hip.dposv_vbatched(ns, pa[i], ns, px[i], info, n, handle=handle), restore,
                          copies, steps, warmup, stream)
    fwd = float(np.mean(ft))
    ok_f = bool(torch.all(info == 0).item())
    for i in range(copies):
        restore(i)
        hip.dposv_vbatched(ns, pa[i], ns, px[i], info, n, handle=handle)
    bt, _ = rotated_calls(lambda i: hip.dposv_backward_vbatched(ns, pa[i], ns, px[i], pw, pab, ns, pbb, info, n,
                                                                handle=handle), None, copies, steps, warmup, stream)
    bwd = float(np.mean(bt))
    ok_b = bool(torch.all(info == 0).item())
    # the Cholesky adjoint: Abar from each copy's L and a lower-triangular Lbar
    lbar = torch.tril(torch.randn((count, n, n), dtype=torch.float64, device=dev, generator=g))
    plb = ptrs(lbar, n * n)
    ct, _ = rotated_calls(lambda i: hip.dpotrf_backward_vbatched(ns, pa[i], ns, plb, ns, pab, ns, info, n,
                                                                 handle=handle), None, copies, steps, warmup, stream)
    cbwd = float(np.mean(ct))
    ok_c = bool(torch.all(info == 0).item())
    handle.close()
    fb = count * 8.0 * (n * (n + 1) / 2 + n * n + 2 * n)
    bb = count * 8.0 * (2 * n * n + 3 * n)
    cb = count * 8.0 * (n * (n + 1) + n * n)
    res = {"workload": f"{count} SPD solves n = {n} (exact mode) and their adjoint (dL/dA, dL/db for L = w.x)",
           "working_copies": copies, "forward_ms": round(fwd, 5), "backward_ms": round(bwd, 5),
           "forward_plus_backward_per_s": round(count / ((fwd + bwd) / 1e3), 1), "all_info_zero": ok_f and ok_b and ok_c,
           "forward_roofline": hbm_entry(fb, fwd), "backward_roofline": hbm_entry(bb, bwd),
           "cholesky_backward": {"workload": f"{count} Cholesky adjoints n = {n}: dloss/dA from L and dloss/dL",
                                 "ms_per_call": round(cbwd, 5), "per_s": round(count / (cbwd / 1e3), 1),
                                 "roofline": hbm_entry(cb, cbwd)}}
    if cpu:
        import oracle

        threads = host_cpu()["threads"]
        a_host = src.cpu().numpy().reshape(-1)
        b_host = b0.cpu().numpy().reshape(-1)
        w_host = w.cpu().numpy().reshape(-1)
        nv = np.full(count, n, dtype=np.int32)
        offs = np.arange(count, dtype=np.int64) * n * n
        xoffs = np.arange(count, dtype=np.int64) * n

        def run():
            aa, xx = a_host.copy(), b_host.copy()
            t0 = time.perf_counter()
            info_c = oracle.dposv_batch(nv, nv, offs, aa, xoffs, xx, threads=threads)
            oracle.dposv_backward_batch(count, n, n, n * n, aa, xx, w_host, threads=threads)
            dt = time.perf_counter() - t0
            assert np.all(info_c == 0)
            return dt

        l_host = a[0].cpu().numpy().reshape(-1)  # copy 0's factors (the posv left L in place)
        lb_host = lbar.cpu().numpy().reshape(-1)

        def run_chol():
            ab = np.empty_like(l_host)
            t0 = time.perf_counter()
            oracle.dpotrf_backward_batch(count, n, n, n * n, l_host, lb_host, ab, threads=threads)
            return time.perf_counter() - t0

        cmed, creps = timed_cpu(run_chol, 4.0, min_reps=3, max_reps=10)
        res["cholesky_backward"]["cpu_baseline"] = cpu_entry(
            count / cmed, "adjoints/s", f"the full {count}-matrix batch, the oracle's Cholesky adjoint "
            f"(oracle/matflat_oracle.c), OpenMP dynamic over {threads} threads, median of {creps} runs", creps)
        med, reps = timed_cpu(run, 4.0, min_reps=3, max_reps=10)
        res["cpu_baseline"] = cpu_entry(count / med, "forward+backward/s",
                                        f"the full {count}-matrix batch: Cholesky + two solves in MatFlat's order, "
                                        "then the adjoint's two solves and symmetric outer product "
                                        f"(oracle/matflat_oracle.c), OpenMP dynamic over {threads} threads, "
                                        f"median of {reps} runs", reps)
    log(f"  adjoint n={n} x{count}: forward {fwd:.4f} ms ({res['forward_roofline']['achieved']} GB/s), backward "
        f"{bwd:.4f} ms ({res['backward_roofline']['achieved']} GB/s); cholesky backward {cbwd:.4f} ms "
        f"({res['cholesky_backward']['roofline']['achieved']} GB/s)"
        + (f"; cpu {res['cpu_baseline']['value']:.0f}/s" if cpu else ""))
    return res


def gemm_leg(steps: int, warmup: int, count: int = 200_000, seed: int = 5, complex_: bool = False,
             cpu: bool = True, sample: int = 20_000):
    """C = alpha op(A) op(B) + beta C over `count` products with m, n, k iid
    U{8..128}, op uniform over {N, T} ({N, T, C} for complex), leading
    dimensions padded to a multiple of 4, alpha, beta ~ U(-1, 1), exact mode
    (MatFlat's summation order)."""
    import torch

    from matflat_amd import hip

    rng = np.random.default_rng(seed + (100 if complex_ else 0))
    m, n, k = (rng.integers(8, 129, size=count).astype(np.int64) for _ in range(3))
    ops = np.array([111, 112, 113]) if complex_ else np.array([111, 112])
    ta, tb = rng.choice(ops, size=count), rng.choice(ops, size=count)
    ra = np.where(ta == 111, m, k)
    rb = np.where(tb == 111, k, n)
    lda, ldb, ldc = (ra + 3) // 4 * 4, (rb + 3) // 4 * 4, (m + 3) // 4 * 4
    size_a = lda * np.where(ta == 111, k, m)
    size_b = ldb * np.where(tb == 111, n, k)
    size_c = ldc * n
    w = 2 if complex_ else 1
    dev = torch.device("cuda")

    def slab(sizes):
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]]) * w
        g = torch.Generator(device=dev)
        g.manual_seed(int(rng.integers(1 << 30)))
        t = torch.randn(int(sizes.sum()) * w, dtype=torch.float64, device=dev, generator=g)
        return t, (t.data_ptr() + 8 * torch.from_numpy(off).to(dev)).contiguous(), off

    A, pa, oa = slab(size_a)
    Bt, pb, ob = slab(size_b)
    C, pc, oc = slab(size_c)
    alpha = torch.from_numpy(rng.uniform(-1, 1, size=count * w)).to(dev)
    beta = torch.from_numpy(rng.uniform(-1, 1, size=count * w)).to(dev)

    def i32(v):
        return torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).to(dev)

    args = [i32(ta), i32(tb), i32(m), i32(n), i32(k), alpha, pa, i32(lda), pb, i32(ldb), beta, pc, i32(ldc)]
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    mf, nf, kf = m.astype(np.float64), n.astype(np.float64), k.astype(np.float64)
    flops = (8.0 if complex_ else 2.0) * mf * nf * kf
    alg = 8.0 * w * (mf * kf + kf * nf + 2 * mf * nf)
    # per product: the slower of its FP64 flops at 78.6 TFLOP/s and its bytes at 8 TB/s
    ceiling_ms = float(np.sum(np.maximum(flops / 78.6e12, alg / 8e12))) * 1e3
    name = "complex128" if complex_ else "fp64"
    res = {"workload": f"{count} {name} products C = alpha op(A) op(B) + beta C, m, n, k ~ U{{8..128}}, "
                       f"op in {{{'N, T, C' if complex_ else 'N, T'}}}, ld padded to 4, alpha, beta ~ U(-1, 1)",
           "model_ceiling_ms": round(ceiling_ms, 4), "alg_bytes": float(alg.sum()), "flops": float(flops.sum()),
           "note": "C is overwritten by each call, so later calls read the previous result (the same work)"}
    if cpu:
        # the first `sample` products from the pristine inputs: the oracle's time, and the bits of
        # one exact-mode GPU call on the same inputs
        import oracle

        threads = host_cpu()["threads"]
        S = min(sample, count)
        ea, eb, ec = int(oa[S - 1] + size_a[S - 1] * w), int(ob[S - 1] + size_b[S - 1] * w), \
            int(oc[S - 1] + size_c[S - 1] * w)
        ha, hb, hc = A[:ea].cpu().numpy(), Bt[:eb].cpu().numpy(), C[:ec].cpu().numpy()
        al_h, be_h = alpha.cpu().numpy()[:S * w], beta.cpu().numpy()[:S * w]
        handle.set_math_mode(hip.MATH_EXACT)
        hip.gemm_vbatched(*args, info, 128, 128, 128, complex_=complex_, handle=handle)
        gpu_c = C[:ec].cpu().numpy()
        out = {}

        def run():
            cc = hc.copy()
            t0 = time.perf_counter()
            inf_c = oracle.gemm_vbatch(int(complex_), ta[:S], tb[:S], m[:S], n[:S], k[:S], al_h, ha, oa[:S],
                                       lda[:S], hb, ob[:S], ldb[:S], be_h, cc, oc[:S], ldc[:S], threads=threads)
            dt = time.perf_counter() - t0
            assert np.all(inf_c == 0)
            out["c"] = cc
            return dt

        med, reps = timed_cpu(run, 4.0, min_reps=3, max_reps=5)
        res["cpu_baseline"] = cpu_entry(S / med, "products/s",
                                        f"the first {S} of the {count} products (the batch's own shape mix), "
                                        "Blas.MulMatMat in MatFlat's order (oracle/matflat_oracle.c), OpenMP dynamic "
                                        f"over {threads} threads, median of {reps} runs", reps)
        res["output_check"] = {"exact_mode_bit_exact_on_sample": bool(np.array_equal(gpu_c, out["c"])),
                               "sample": S}
        log(f"  gemm {name}: cpu {res['cpu_baseline']['value']:.0f} products/s; exact-mode bits on the sample "
            f"{'match' if res['output_check']['exact_mode_bit_exact_on_sample'] else 'DIFFER'}")
    for mode_name, mode in (("fast", hip.MATH_FAST), ("exact", hip.MATH_EXACT)):
        handle.set_math_mode(mode)
        ms = timed_calls(lambda: hip.gemm_vbatched(*args, info, 128, 128, 128, complex_=complex_, handle=handle),
                         None, steps, warmup, stream)
        ok = bool(torch.all(info == 0).item())
        # the same batch as one captured graph (tile-list build + GEMM kernel), replayed
        plan = hip.GemmPlan(*args, info, 128, 128, 128, complex_=complex_, handle=handle)
        ms_graph = timed_calls(plan.execute, None, steps, warmup, stream)
        plan.close()
        ok = ok and bool(torch.all(info == 0).item())
        res[mode_name] = {"ms_per_call": round(ms, 5), "ms_per_graph_replay": round(ms_graph, 5),
                          "products_per_s": round(count / (ms / 1e3), 1),
                          "tflops": round(float(flops.sum()) / (ms / 1e3) / 1e12, 2), "all_info_zero": ok,
                          "roofline": hbm_entry(float(alg.sum()), ms),
                          "frac_of_model_ceiling": round(ceiling_ms / ms, 4),
                          "note": ("FP64 matrix cores, within 4 k eps of MatFlat" if mode == hip.MATH_FAST
                                   else "MatFlat's summation order, bit-identical")}
        log(f"  gemm {name} {mode_name} x{count}: {ms:.4f} ms/call ({ms_graph:.4f} as a graph), "
            f"{res[mode_name]['tflops']} TFLOP/s, "
            f"ceiling {ceiling_ms:.3f} ms ({res[mode_name]['frac_of_model_ceiling']:.3f})")
    handle.close()
    del A, Bt, C
    torch.cuda.empty_cache()
    return res


def trsv_leg(steps: int, warmup: int, count: int = 100_000, seed: int = 7):
    """Blas.SolveTriangular in its four non-unit forms over `count` triangles
    with n ~ U{10..100}, lda = 8 ceil(n / 8), N(0,1) entries and a diagonal
    of n + |N(0,1)|, unit stride, exact mode (bit-identical).  HBM bound:
    8 (n (n + 1) / 2 + 2 n) + 4 bytes per solve."""
    import torch

    from matflat_amd import hip

    rng = np.random.default_rng(seed)
    ns = rng.integers(10, 101, size=count).astype(np.int64)
    ldas = (ns + 7) // 8 * 8
    sizes = ldas * ns
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    xoffs = np.concatenate([[0], np.cumsum(ns)[:-1]])
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A = torch.randn(int(sizes.sum()), dtype=torch.float64, device=dev, generator=g)
    rep = np.repeat(np.arange(count), ns)
    jj = np.arange(int(ns.sum())) - np.repeat(xoffs, ns)
    didx = torch.from_numpy(offs[rep] + jj * (ldas[rep] + 1)).to(dev)
    A[didx] = A[didx].abs() + torch.from_numpy(ns[rep].astype(np.float64)).to(dev)
    x0 = torch.randn(int(ns.sum()), dtype=torch.float64, device=dev, generator=g)
    x = x0.clone()
    a_ptrs = (A.data_ptr() + 8 * torch.from_numpy(offs).to(dev)).contiguous()
    x_ptrs = (x.data_ptr() + 8 * torch.from_numpy(xoffs).to(dev)).contiguous()
    n_t = torch.from_numpy(ns.astype(np.int32)).to(dev)
    lda_t = torch.from_numpy(ldas.astype(np.int32)).to(dev)
    inc = torch.ones(count, dtype=torch.int32, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    alg = float(np.sum(8.0 * (ns * (ns + 1) / 2 + 2 * ns) + 4))
    res = {"workload": f"{count} triangular solves per form, n ~ U{{10..100}}, lda = 8 ceil(n/8), unit stride, "
                       "exact mode (MatFlat's order, bit-identical)", "alg_bytes": alg}
    for name, uplo, trans in (("lower_notrans", hip.LOWER, hip.NOTRANS), ("lower_trans", hip.LOWER, hip.TRANS),
                              ("upper_notrans", hip.UPPER, hip.NOTRANS), ("upper_trans", hip.UPPER, hip.TRANS)):
        ms = timed_calls(lambda: hip.dtrsv_vbatched(uplo, trans, hip.NONUNIT, n_t, a_ptrs, lda_t, x_ptrs, inc, info,
                                                    100, handle=handle), lambda: x.copy_(x0), steps, warmup, stream)
        res[name] = {"ms_per_call": round(ms, 5), "solves_per_s": round(count / (ms / 1e3), 1),
                     "all_info_zero": bool(torch.all(info == 0).item()), "roofline": hbm_entry(alg, ms)}
        log(f"  trsv {name} x{count}: {ms:.4f} ms/call, {res[name]['roofline']['achieved']} GB/s")
    handle.close()
    del A, x, x0
    torch.cuda.empty_cache()
    return res


def gemv_leg(steps: int, warmup: int, count: int = 200_000, seed: int = 6):
    """y = op(A) x (Blas.MulMatVec) over `count` products with m, n iid
    U{8..128}, op uniform over {N, T}, lda padded to a multiple of 4, unit
    strides, MatFlat's summation order (bit-identical).  HBM bound: the
    algorithmic bytes are 8 (m n + len(x) + len(y)) per product."""
    import torch

    from matflat_amd import hip

    rng = np.random.default_rng(seed)
    m, n = (rng.integers(8, 129, size=count).astype(np.int64) for _ in range(2))
    tr = rng.choice(np.array([111, 112]), size=count)
    lda = (m + 3) // 4 * 4
    xl, yl = np.where(tr == 112, m, n), np.where(tr == 112, n, m)
    dev = torch.device("cuda")

    def slab(sizes):
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        g = torch.Generator(device=dev)
        g.manual_seed(int(rng.integers(1 << 30)))
        t = torch.randn(int(sizes.sum()), dtype=torch.float64, device=dev, generator=g)
        return t, (t.data_ptr() + 8 * torch.from_numpy(off).to(dev)).contiguous()

    A, pa = slab(lda * n)
    X, px = slab(xl)
    Y, py = slab(yl)

    def i32(v):
        return torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).to(dev)

    ones = i32(np.ones(count))
    args = [i32(tr), i32(m), i32(n), pa, i32(lda), px, ones, py, ones]
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    alg = float(np.sum(8.0 * (m * n + xl + yl)))
    ms = timed_calls(lambda: hip.dgemv_vbatched(*args, info, 128, 128, handle=handle), None, steps, warmup, stream)
    ok = bool(torch.all(info == 0).item())
    res = {"workload": f"{count} fp64 products y = op(A) x, m, n ~ U{{8..128}}, op in {{N, T}}, lda padded to 4, "
                       "unit strides, exact mode (MatFlat's summation order, bit-identical)",
           "ms_per_call": round(ms, 5), "products_per_s": round(count / (ms / 1e3), 1), "all_info_zero": ok,
           "roofline": hbm_entry(alg, ms)}
    log(f"  gemv fp64 x{count}: {ms:.4f} ms/call, {res['roofline']['achieved']} GB/s")
    handle.close()
    del A, X, Y
    torch.cuda.empty_cache()
    return res


def gevd_leg(steps: int, warmup: int, count: int = 10_000, nmin: int = 8, nmax: int = 64, seed: int = 8,
             cpu: bool = True):
    """Factorization.Gevd (double, MatFlat's order) over `count` pairs, n ~
    U{nmin..nmax}: A = X X^T, B = Y Y^T + n I / 4 with X, Y ~ N(0, 1), upper
    triangles stored (lower triangles NaN, never read), lda = ldb = n.  The call
    overwrites a and b, so both are restored (untimed) before every call."""
    import torch

    from matflat_amd import hip

    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count).astype(np.int32)
    sizes = ns.astype(np.int64) ** 2
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    a_host = np.empty(int(sizes.sum()))
    b_host = np.empty(int(sizes.sum()))
    for i, n in enumerate(ns):
        x = rng.standard_normal((n, n))
        y = rng.standard_normal((n, n))
        am = x @ x.T
        bm = y @ y.T + (n / 4.0) * np.eye(n)
        low = np.tril_indices(n, -1)
        am[low] = np.nan
        bm[low] = np.nan
        a_host[offs[i]:offs[i] + n * n] = am.T.reshape(-1)  # column-major
        b_host[offs[i]:offs[i] + n * n] = bm.T.reshape(-1)
    woffs = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]]).astype(np.int64)
    dev = torch.device("cuda")
    a_src = torch.from_numpy(a_host).to(dev)
    b_src = torch.from_numpy(b_host).to(dev)
    a, b = a_src.clone(), b_src.clone()
    w = torch.zeros(int(ns.sum()), dtype=torch.float64, device=dev)
    o = torch.from_numpy(offs).to(dev)
    pa = (a.data_ptr() + 8 * o).contiguous()
    pb = (b.data_ptr() + 8 * o).contiguous()
    pw = (w.data_ptr() + 8 * torch.from_numpy(woffs).to(dev)).contiguous()
    n_t = torch.from_numpy(ns).to(dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)

    def restore():
        a.copy_(a_src)
        b.copy_(b_src)

    ms = timed_calls(lambda: hip.dgevd_vbatched(n_t, pa, n_t, pb, n_t, pw, info, nmax, handle=handle), restore,
                     steps, warmup, stream)
    torch.cuda.synchronize()
    ok = bool(torch.all(info == 0).item())
    # one sampled matrix per size range checked against the oracle, bit for bit
    check = {}
    if cpu:
        import oracle

        ga, gb, gw = a.cpu().numpy(), b.cpu().numpy(), w.cpu().numpy()
        same = True
        for i in list(range(0, count, max(1, count // 16))):
            n = int(ns[i])
            aa = a_host[offs[i]:offs[i] + n * n].copy()
            bb = b_host[offs[i]:offs[i] + n * n].copy()
            ww = np.zeros(n)
            oracle.dgevd(n, aa, n, bb, n, ww)
            same &= bool(np.array_equal(aa.view(np.uint64), ga[offs[i]:offs[i] + n * n].view(np.uint64))
                         and np.array_equal(bb.view(np.uint64), gb[offs[i]:offs[i] + n * n].view(np.uint64))
                         and np.array_equal(ww.view(np.uint64), gw[woffs[i]:woffs[i] + n].view(np.uint64)))
        check = {"sampled": len(range(0, count, max(1, count // 16))), "bit_exact": same}
    handle.close()
    n64 = ns.astype(np.float64)
    alg = float(np.sum(8.0 * (n64 * (n64 + 1) + 2 * n64 * n64 + n64) + 4))
    res = {"workload": f"{count} generalized eigenproblems A v = w B v (Factorization.Gevd, MatFlat's order), "
                       f"n ~ U{{{nmin}..{nmax}}}, A = X X^T, B = Y Y^T + n I / 4, X, Y ~ N(0, 1)",
           "ms_per_call": round(ms, 4), "problems_per_s": round(count / (ms / 1e3), 1), "all_info_zero": ok,
           "roofline": hbm_entry(alg, ms), "output_check": check,
           "note": "one wavefront per problem; the Householder, QR-sweep and rotation chains are sequential, so "
                   "the time is latency, not bytes"}
    if cpu:
        import oracle

        threads = host_cpu()["threads"]

        def run():
            aa, bb = a_host.copy(), b_host.copy()
            ww = np.zeros(int(ns.sum()))
            t0 = time.perf_counter()
            info_c = oracle.dgevd_batch(ns, ns, ns, offs, offs, woffs, aa, bb, ww, threads=threads)
            dt = time.perf_counter() - t0
            assert np.all(info_c == 0)
            return dt

        med, reps = timed_cpu(run, 10.0, min_reps=3, max_reps=5)
        res["cpu_baseline"] = cpu_entry(count / med, "problems/s",
                                        f"the full {count}-pair batch, Factorization.Gevd in MatFlat's order "
                                        f"(oracle/matflat_oracle.c), OpenMP dynamic over {threads} threads, "
                                        f"median of {reps} runs", reps)
    log(f"  gevd x{count} n~U{{{nmin}..{nmax}}}: {ms:.3f} ms/call, {res['problems_per_s']:.0f}/s, "
        f"{res['roofline']['achieved']} GB/s; check {check}"
        + (f"; cpu {res['cpu_baseline']['value']:.0f}/s" if cpu else ""))
    del a, b, a_src, b_src
    torch.cuda.empty_cache()
    return res


LEGS = ("lu", "adjoint", "gemm", "zgemm", "gemv", "trsv", "gevd")


def other_legs(steps: int, cpu: bool = True, which=LEGS) -> dict:
    log("other configurations")
    run = {"lu": ("lu_n64", lambda: lu_leg(steps, 2, cpu=cpu)),
           "adjoint": ("spd_solve_adjoint_n32", lambda: adjoint_leg(steps, 2, cpu=cpu)),
           "gemm": ("gemm_fp64", lambda: gemm_leg(max(3, steps // 4), 1, cpu=cpu)),
           "zgemm": ("gemm_complex128", lambda: gemm_leg(max(3, steps // 4), 1, complex_=True, cpu=cpu)),
           "gemv": ("gemv_fp64", lambda: gemv_leg(steps, 2)),
           "trsv": ("trsv", lambda: trsv_leg(steps, 2)),
           "gevd": ("gevd", lambda: gevd_leg(max(3, steps // 4), 1, cpu=cpu))}
    out = {}
    for w in which:
        key, fn = run[w]
        out[key] = fn()
    return out


def cpu_baseline(buf_host, offsets, ns, ldas, xoff, b_host, max_seconds=20.0):
    """The oracle's Cholesky + two solves (MatFlat's operation order) over the
    whole batch on the host's cores: OpenMP dynamic scheduling, one matrix per
    iteration."""
    import oracle

    threads = host_cpu()["threads"]
    count = ns.size

    def run():
        a = buf_host.copy()
        x = b_host.copy()
        t0 = time.perf_counter()
        info = oracle.dposv_batch(ns, ldas, offsets, a, xoff, x, threads=threads)
        dt = time.perf_counter() - t0
        assert np.all(info == 0)
        return dt

    med, reps = timed_cpu(run, max_seconds, min_reps=3, max_reps=5)
    return cpu_entry(count / med, "solves/s",
                     f"the full {count}-matrix batch, C restatement of MatFlat (oracle/matflat_oracle.c, -O2 "
                     f"-ffp-contract=off) with OpenMP dynamic scheduling over {threads} threads, median of {reps} runs",
                     reps)


class SpdCopies:
    """Working copies of the SPD batch for in-place solves: `copies` buffers
    of the matrices and right-hand sides with their pointer arrays and info
    vectors, restored from the pristine batch by restore(i)."""

    def __init__(self, buf, b, offsets, ns, ldas, xoff, copies):
        import torch

        dev = buf.device
        self.src, self.b = buf, b
        self.copies = copies
        self.n_t = torch.from_numpy(ns).to(dev)
        self.lda_t = torch.from_numpy(ldas).to(dev)
        offs = 8 * torch.from_numpy(offsets).to(dev)
        xoffs = 8 * torch.from_numpy(xoff).to(dev)
        self.work = [torch.empty_like(buf) for _ in range(copies)]
        self.x = [torch.empty_like(b) for _ in range(copies)]
        self.a_ptrs = [(w.data_ptr() + offs).contiguous() for w in self.work]
        self.x_ptrs = [(x.data_ptr() + xoffs).contiguous() for x in self.x]
        self.info = [torch.zeros(ns.size, dtype=torch.int32, device=dev) for _ in range(copies)]

    def restore(self, i):
        self.work[i].copy_(self.src)
        self.x[i].copy_(self.b)
        self.info[i].fill_(-99)


def check_spd_outputs(batch, i, offsets, ns, ldas, xoff, sample=256, seed=11):
    """Checks the outputs a solve left in working copy i: every info zero, and
    for a seeded sample of matrices the factor residual ||L L^T - A||_1 /
    (n ||A||_1 eps) <= 10, an exactly zero strict upper triangle and a solve
    backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= n
    eps, with A rebuilt from the pristine batch."""
    eps = np.finfo(np.float64).eps
    info = batch.info[i].cpu().numpy()
    rng = np.random.default_rng(seed)
    pick = rng.choice(ns.size, size=min(sample, ns.size), replace=False)
    worst_f, worst_b, upper_ok = 0.0, 0.0, True
    for j in pick:
        n, lda, o, xo = int(ns[j]), int(ldas[j]), int(offsets[j]), int(xoff[j])
        a = batch.src[o:o + n * lda].cpu().numpy().reshape(n, lda)[:, :n].T  # a[r, c], upper triangle valid
        a = np.triu(a) + np.triu(a, 1).T
        lf = batch.work[i][o:o + n * lda].cpu().numpy().reshape(n, lda)[:, :n].T
        upper_ok = upper_ok and bool(np.all(np.triu(lf, 1) == 0.0))
        lo = np.tril(lf)
        worst_f = max(worst_f, float(np.linalg.norm(lo @ lo.T - a, 1) / (n * np.linalg.norm(a, 1) * eps)))
        x = batch.x[i][xo:xo + n].cpu().numpy()
        b = batch.b[xo:xo + n].cpu().numpy()
        r = b - a @ x
        be = float(np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max()))
        worst_b = max(worst_b, float(be / (n * eps)))
    res = {"info_all_zero": bool(np.all(info == 0)), "sampled": int(pick.size),
           "max_factor_residual_ratio": round(worst_f, 4), "max_backward_error_over_n_eps": round(worst_b, 4),
           "strict_upper_zero": upper_ok}
    res["ok"] = bool(res["info_all_zero"] and upper_ok and worst_f <= 10.0 and worst_b <= 1.0)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copies", type=int, default=2,
                    help="working copies of the headline batch the timed steps rotate over (at least 2)")
    ap.add_argument("--math", choices=["fast", "exact"], default="fast",
                    help="math mode of the headline run (the other mode is timed as a secondary line)")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision LU-solve comparison (n = 128..512)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the LU (n = 64), adjoint (n = 32) and GEMM legs")
    ap.add_argument("--only", choices=["mixed", "legs"] + list(LEGS), default=None,
                    help="run only the mixed-precision comparison or only the other-configuration legs and print "
                         "their JSON (for profiling them alone); the default runs everything")
    ap.add_argument("--save-mixed", default=None, metavar="DIR",
                    help="generate the mixed-precision batch, write it to DIR and exit")
    ap.add_argument("--load-mixed", default=None, metavar="DIR",
                    help="with --only mixed: read the batch --save-mixed wrote instead of generating it")
    ap.add_argument("--bucket-streams", type=int, default=0,
                    help="streams the FAST size-bucket launches are forked over (0 = the library default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device copy (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 (matrices are independent: "
                                   "a multi-GPU run would give each GPU a slice of the size-sorted batch)"}))
        return 2

    import torch

    from matflat_amd import hip

    cpu_on = not args.no_cpu_baseline
    if args.only is not None and args.calibrate_copy:
        copy_calibration()
    if args.save_mixed:
        save_conditioned_batch(args.save_mixed)
        return 0
    if args.only == "mixed":
        print(json.dumps({"mixed_precision": mixed_leg(args.steps, args.warmup, cpu=cpu_on,
                                                       batch_dir=args.load_mixed)}, default=_plain), flush=True)
        return 0
    if args.only is not None:
        which = LEGS if args.only == "legs" else (args.only,)
        print(json.dumps({"other_configs": other_legs(args.steps, cpu=cpu_on, which=which)}, default=_plain),
              flush=True)
        return 0

    seed, nmin, nmax = 2, 10, 100
    log(f"generating {args.count} SPD matrices n~U{{{nmin}..{nmax}}}")
    buf, offsets, ns, ldas, xoff, b = make_spd_batch(args.count, nmin, nmax, seed)
    torch.cuda.synchronize()
    buf_host = buf.cpu().numpy() if not args.no_cpu_baseline else None
    b_host = b.cpu().numpy() if not args.no_cpu_baseline else None

    copies = max(2, args.copies)
    batch = SpdCopies(buf, b, offsets, ns, ldas, xoff, copies)
    log(f"{copies} working copies of {buf.numel() * 8 / 1e9:.2f} GB")
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    if args.bucket_streams:
        handle.set_bucket_streams(args.bucket_streams)

    def run_mode(mode: int, steps: int, warmup: int):
        """Per-step ms of plan replays on rotated pristine copies (one plan per
        copy) and the copy the last timed step used."""
        handle.set_math_mode(mode)
        plans = [hip.DposvPlan(batch.n_t, batch.a_ptrs[i], batch.lda_t, batch.x_ptrs[i], batch.info[i], nmax,
                               handle=handle) for i in range(copies)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        times, last = rotated_calls(lambda i: plans[i].execute(), batch.restore, copies, steps, warmup, stream)
        wall = time.perf_counter() - t0
        for p in plans:
            p.close()
        return times, last, wall * 1e3 / (steps + warmup)

    main_mode = hip.MATH_FAST if args.math == "fast" else hip.MATH_EXACT
    times, last, wall_ms = run_mode(main_mode, args.steps, args.warmup)
    ms = float(np.mean(times))
    check = check_spd_outputs(batch, last, offsets, ns, ldas, xoff)
    log(f"{args.math}: {args.steps} steps: {ms:.4f} ms/step mean, {float(np.median(times)):.4f} median (events); "
        f"wall incl. restores {wall_ms:.3f} ms/step; output check {check}")
    if not check["ok"]:
        raise RuntimeError(f"the outputs of the last timed step failed their check: {check}")
    secondary = None
    if not args.no_secondary:
        other = hip.MATH_EXACT if main_mode == hip.MATH_FAST else hip.MATH_FAST
        otimes, olast, _ = run_mode(other, max(4, args.steps // 4), 2)
        oms = float(np.mean(otimes))
        ocheck = check_spd_outputs(batch, olast, offsets, ns, ldas, xoff, sample=64)
        oname = "exact" if other == hip.MATH_EXACT else "fast"
        log(f"{oname}: {oms:.4f} ms/step; output check {ocheck}")
        secondary = {"math": oname, "value": round(args.count / (oms / 1e3), 1), "unit": "solves/s",
                     "ms_per_step": round(oms, 5), "steps": len(otimes), "output_check": ocheck,
                     "note": ("bit-identical to MatFlat's CPU results" if oname == "exact"
                              else "FMA / right-looking, backward error O(n eps)")}

    value = args.count / (ms / 1e3)
    log("kernel timing pass")
    kern = kernel_pass(handle, main_mode, batch, nmax, ns, args.steps, stream)
    roofline = roofline_of(kern, ns, default_workload=args.count == 100_000)
    handle.close()
    if args.calibrate_copy:
        copy_calibration()
    del batch
    torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(buf_host, offsets, ns, ldas, xoff, b_host)
        log(f"  cpu: {cpu['value']:.0f} solves/s on {cpu['cores']} threads ({cpu['cpu_model']})")
    del buf_host, b_host
    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(max(3, args.steps // 4), 1, cpu=not args.no_cpu_baseline)
    legs = None
    if not args.no_legs:
        del buf, b
        torch.cuda.empty_cache()
        legs = other_legs(args.steps, cpu=not args.no_cpu_baseline)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1); A = B B^T + n I)",
        "config": {"workload": "batched FP64 Cholesky + forward/back triangular solve, "
                               f"{args.count} SPD matrices, n~U{{{nmin}..{nmax}}}, lda=8*ceil(n/8), one rhs each, "
                               "one hipGraph replay per step on a pristine working copy",
                   "count": args.count, "n_min": nmin, "n_max": nmax, "seed": seed, "working_copies": copies},
        "ms_per_step_median": round(float(np.median(times)), 5),
        "output_check": check,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "math": args.math,
        "secondary": secondary,
        "mixed_precision": mixed,
        "other_configs": legs,
    }
    print(json.dumps(out, default=_plain), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
