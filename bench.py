"""Benchmark: variable-size batched FP64 SPD solve (Cholesky + two triangular
solves) on one MI355X.

Workload (one step): 100,000 SPD matrices, order n ~ U{10..100}, leading
dimension 8*ceil(n/8), A = B B^T + n I with B ~ N(0,1) (upper triangle
stored, strict lower triangle NaN: it is never read), one N(0,1) right-hand
side each.  A step is one replay of the hipGraph that holds the whole batch
(device bucketing + the persistent solve kernels); inputs are resident in
HBM before timing starts.  Each step factors and solves in place; the
kernels are data-independent (no pivoting, no early exits on SPD input), so
every step performs the same work on the same bytes.

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


def kernel_pass(handle, mode, n_t, a_ptrs, lda_t, x_ptrs, info, nmax, ns, steps):
    """Times every kernel launch of `steps` eager batched solves with HIP
    events recorded by the library on the stream it launches on.  Returns
    one entry per distinct launch: kind, row bound, matrices it processes,
    their algorithmic bytes and the mean duration."""
    import torch

    from matflat_amd import _lib, hip

    handle.set_math_mode(mode)
    handle.set_timing(True)
    torch.cuda.synchronize()
    for _ in range(steps):
        hip.dposv_vbatched(n_t, a_ptrs, lda_t, x_ptrs, info, nmax, handle=handle)
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


def mixed_leg(steps: int, warmup: int, count: int = 8000, seed: int = 3):
    """FP32/BF16-factor + FP64-refinement LU solve against the plain FP64
    blocked LU solve on the same batch, both to backward error <= 1e-15 n."""
    import torch

    from matflat_amd import _lib, hip

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
    flops = float(np.sum(2.0 / 3.0 * ns.astype(np.float64) ** 3 + 1.5 * ns.astype(np.float64) ** 2))
    for name, mixed in (("fp64", False), ("mixed", True)):
        for _ in range(warmup):
            call(mixed)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            call(mixed)
        stop.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
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
        log(f"  {name}: {ms:.3f} ms/step, {entry['solves_per_s']:.0f} solves/s, {entry['hpl_tflops']} TFLOP/s (HPL "
            f"count), max berr/n {entry['max_berr_over_n']:.2e}" +
            (f", iters mean {entry['mean_iters']} max {entry['max_iters']}" if mixed else ""))
        res[name] = entry
    handle.close()
    res["speedup_mixed_over_fp64"] = round(res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"], 3)
    res["workload"] = (f"{count} LU solves, n uniform over {list(MIXED_SIZES)}, A = U diag(s) V^T (Haar U, V), "
                       "cond 10^U(2,6), b ~ N(0,1); refine to ||b-Ax||/(||A|| ||x|| + ||b||) <= 1e-15 n or 10 "
                       "iterations")
    res["seed"] = seed
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


def lu_leg(steps: int, warmup: int, count: int = 1024, n: int = 64, seed: int = 1):
    """Exact (bit-identical to MatFlat) partial-pivot LU of `count` column-major
    n x n matrices, lda = n, N(0,1) entries; each timed call factors a fresh
    copy of the batch."""
    import torch

    from matflat_amd import hip

    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    src = torch.randn(count * n * n, dtype=torch.float64, device=dev, generator=g)
    a = torch.empty_like(src)
    piv = torch.zeros(count * n, dtype=torch.int32, device=dev)
    sign = torch.zeros(count, dtype=torch.int32, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    handle.set_math_mode(hip.MATH_EXACT)
    ms = timed_calls(lambda: hip.dgetrf_strided_batched(count, n, n, a, n, n * n, piv, n, sign, info, handle=handle),
                     lambda: a.copy_(src), steps, warmup, stream)
    torch.cuda.synchronize()
    ok = bool(torch.all(info == 0).item())
    # Factorization.LuInverse of the factored batch; each timed call restores the factors
    lu_src = a.clone()
    idx = torch.arange(count, dtype=torch.int64, device=dev)
    a_ptrs = a.data_ptr() + 8 * n * n * idx
    piv_ptrs = piv.data_ptr() + 4 * n * idx
    n_t = torch.full((count,), n, dtype=torch.int32, device=dev)
    inv_info = torch.zeros(count, dtype=torch.int32, device=dev)
    ms_inv = timed_calls(lambda: hip.dgetri_vbatched(n_t, a_ptrs, n_t, piv_ptrs, inv_info, n, handle=handle),
                         lambda: a.copy_(lu_src), steps, warmup, stream)
    inv_ok = bool(torch.all(inv_info == 0).item())
    # LU solve (the LuInverse composition for one right-hand side) with the factors
    a.copy_(lu_src)
    bvec = torch.randn(count * n, dtype=torch.float64, device=dev, generator=g)
    xvec = torch.zeros_like(bvec)
    b_ptrs = bvec.data_ptr() + 8 * n * idx
    x_ptrs = xvec.data_ptr() + 8 * n * idx
    rs_info = torch.zeros(count, dtype=torch.int32, device=dev)
    ms_rs = timed_calls(lambda: hip.dgetrs_vbatched(n_t, a_ptrs, n_t, piv_ptrs, b_ptrs, x_ptrs, rs_info, n,
                                                    handle=handle), None, steps, warmup, stream)
    rs_ok = bool(torch.all(rs_info == 0).item())
    handle.close()
    alg = count * (16.0 * n * n + 4 * n + 4)
    res = {"workload": f"{count} LU factorizations (partial pivoting), n = {n}, lda = {n}, N(0,1), exact mode "
                       "(bit-identical to MatFlat's Factorization.Lu)",
           "ms_per_call": round(ms, 5), "factorizations_per_s": round(count / (ms / 1e3), 1),
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
    log(f"  LU n={n} x{count}: {ms:.4f} ms/call, {res['factorizations_per_s']:.0f} LU/s, "
        f"{res['roofline']['achieved']} GB/s; inverse {ms_inv:.4f} ms/call; solve {ms_rs:.4f} ms/call")
    return res


def adjoint_leg(steps: int, warmup: int, count: int = 50_000, n: int = 32, seed: int = 4):
    """Forward (exact-mode SPD solve) and hand-written backward of the batched
    SPD solve, n = 32, A = B B^T + n I."""
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
    a, x = torch.empty_like(src), torch.empty_like(b0)
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
    pa, px, pw, pab, pbb = ptrs(a, n * n), ptrs(x, n), ptrs(w, n), ptrs(abar, n * n), ptrs(bbar, n)

    def restore():
        a.copy_(src)
        x.copy_(b0)

    fwd = timed_calls(lambda: hip.dposv_vbatched(ns, pa, ns, px, info, n, handle=handle), restore, steps, warmup,
                      stream)
    ok_f = bool(torch.all(info == 0).item())
    bwd = timed_calls(lambda: hip.dposv_backward_vbatched(ns, pa, ns, px, pw, pab, ns, pbb, info, n, handle=handle),
                      None, steps, warmup, stream)
    ok_b = bool(torch.all(info == 0).item())
    handle.close()
    fb = count * 8.0 * (n * (n + 1) / 2 + n * n + 2 * n)
    bb = count * 8.0 * (2 * n * n + 3 * n)
    res = {"workload": f"{count} SPD solves n = {n} (exact mode) and their adjoint (dL/dA, dL/db for L = w.x)",
           "forward_ms": round(fwd, 5), "backward_ms": round(bwd, 5),
           "forward_plus_backward_per_s": round(count / ((fwd + bwd) / 1e3), 1), "all_info_zero": ok_f and ok_b,
           "forward_roofline": hbm_entry(fb, fwd), "backward_roofline": hbm_entry(bb, bwd)}
    log(f"  adjoint n={n} x{count}: forward {fwd:.4f} ms ({res['forward_roofline']['achieved']} GB/s), backward "
        f"{bwd:.4f} ms ({res['backward_roofline']['achieved']} GB/s)")
    return res


def gemm_leg(steps: int, warmup: int, count: int = 200_000, seed: int = 5, complex_: bool = False):
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
        return t, (t.data_ptr() + 8 * torch.from_numpy(off).to(dev)).contiguous()

    A, pa = slab(size_a)
    Bt, pb = slab(size_b)
    C, pc = slab(size_c)
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
           "model_ceiling_ms": round(ceiling_ms, 4), "alg_bytes": float(alg.sum()), "flops": float(flops.sum())}
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


def other_legs(steps: int) -> dict:
    log("other configurations")
    return {"lu_n64": lu_leg(steps, 2), "spd_solve_adjoint_n32": adjoint_leg(steps, 2),
            "gemm_fp64": gemm_leg(max(3, steps // 4), 1),
            "gemm_complex128": gemm_leg(max(3, steps // 4), 1, complex_=True),
            "gemv_fp64": gemv_leg(steps, 2), "trsv": trsv_leg(steps, 2)}


def cpu_baseline(buf_host, offsets, ns, ldas, xoff, b_host, max_seconds=20.0):
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    count = ns.size
    reps, elapsed = 0, []
    t_start = time.perf_counter()
    while True:
        a = buf_host.copy()
        x = b_host.copy()
        t0 = time.perf_counter()
        info = oracle.dposv_batch(ns, ldas, offsets, a, xoff, x, threads=threads)
        elapsed.append(time.perf_counter() - t0)
        reps += 1
        assert np.all(info == 0)
        if reps >= 5 or time.perf_counter() - t_start > max_seconds:
            break
    med = float(np.median(elapsed))
    return {"value": count / med, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"the full {count}-matrix batch, C restatement of MatFlat (oracle/matflat_oracle.c, "
                      f"-O2 -ffp-contract=off) with OpenMP dynamic scheduling, median of {reps} runs"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--math", choices=["fast", "exact"], default="fast",
                    help="math mode of the headline run (the other mode is timed as a secondary line)")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision LU-solve comparison (n = 128..512)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the LU (n = 64), adjoint (n = 32) and GEMM legs")
    ap.add_argument("--only", choices=["mixed", "legs"], default=None,
                    help="run only the mixed-precision comparison or only the other-configuration legs and print "
                         "their JSON (for profiling them alone); the default runs everything")
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

    if args.only == "mixed":
        print(json.dumps({"mixed_precision": mixed_leg(args.steps, args.warmup)}), flush=True)
        return 0
    if args.only == "legs":
        print(json.dumps({"other_configs": other_legs(args.steps)}), flush=True)
        return 0

    seed, nmin, nmax = 2, 10, 100
    log(f"generating {args.count} SPD matrices n~U{{{nmin}..{nmax}}}")
    buf, offsets, ns, ldas, xoff, b = make_spd_batch(args.count, nmin, nmax, seed)
    torch.cuda.synchronize()
    buf_host = buf.cpu().numpy() if not args.no_cpu_baseline else None
    b_host = b.cpu().numpy() if not args.no_cpu_baseline else None

    dev = torch.device("cuda")
    n_t = torch.from_numpy(ns).to(dev)
    lda_t = torch.from_numpy(ldas).to(dev)
    a_ptrs = (buf.data_ptr() + 8 * torch.from_numpy(offsets).to(dev)).contiguous()
    x_ptrs = (b.data_ptr() + 8 * torch.from_numpy(xoff).to(dev)).contiguous()
    info = torch.zeros(args.count, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    handle = hip.Handle()
    handle.bind_stream(stream)
    if args.bucket_streams:
        handle.set_bucket_streams(args.bucket_streams)

    def run_mode(mode: int, steps: int, warmup: int):
        handle.set_math_mode(mode)
        plan = hip.DposvPlan(n_t, a_ptrs, lda_t, x_ptrs, info, nmax, handle=handle)
        for _ in range(warmup):
            plan.execute()
        torch.cuda.synchronize()
        if int((info != 0).sum().item()) != 0:
            raise RuntimeError("factorization failures during warmup")
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record(stream)
        for _ in range(steps):
            plan.execute()
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        plan.close()
        return start.elapsed_time(stop) / steps, wall * 1e3 / steps

    main_mode = hip.MATH_FAST if args.math == "fast" else hip.MATH_EXACT
    ms, wall_ms = run_mode(main_mode, args.steps, args.warmup)
    log(f"{args.math}: {args.steps} steps: {ms:.4f} ms/step (events), wall {wall_ms:.4f} ms/step")
    secondary = None
    if not args.no_secondary:
        other = hip.MATH_EXACT if main_mode == hip.MATH_FAST else hip.MATH_FAST
        oms, _ = run_mode(other, max(3, args.steps // 4), 1)
        oname = "exact" if other == hip.MATH_EXACT else "fast"
        log(f"{oname}: {oms:.4f} ms/step")
        secondary = {"math": oname, "value": round(args.count / (oms / 1e3), 1), "unit": "solves/s",
                     "ms_per_step": round(oms, 5),
                     "note": ("bit-identical to MatFlat's CPU results" if oname == "exact"
                              else "FMA / right-looking, backward error O(n eps)")}

    value = args.count / (ms / 1e3)
    log("kernel timing pass")
    kern = kernel_pass(handle, main_mode, n_t, a_ptrs, lda_t, x_ptrs, info, nmax, ns, args.steps)
    roofline = roofline_of(kern, ns, default_workload=args.count == 100_000)
    if args.calibrate_copy:
        copy_calibration()
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(buf_host, offsets, ns, ldas, xoff, b_host)
    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(max(3, args.steps // 4), 1)
    legs = None
    if not args.no_legs:
        del buf, b
        torch.cuda.empty_cache()
        legs = other_legs(args.steps)
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
                               "one hipGraph replay per step",
                   "count": args.count, "n_min": nmin, "n_max": nmax, "seed": seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "math": args.math,
        "secondary": secondary,
        "mixed_precision": mixed,
        "other_configs": legs,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
