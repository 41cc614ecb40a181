"""Parity at the full workload sizes, through properties that do not depend
on the batch size (the oracle is far too slow for every matrix):

* 100,000 SPD solves, n ~ U{10..100}, lda = 8 ceil(n/8) (the benchmark's
  workload, built by bench.make_spd_batch): every solve's normwise backward
  error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= n eps, every
  factor's strict upper triangle exactly zero and its padding rows
  untouched, info = 0; in exact mode a random sample of 200 matrices is also
  compared bit for bit with the oracle (MatFlat's Factorization.Cholesky and
  two Blas.SolveTriangular calls).
* 200,000 FP64 products C = alpha op(A) op(B) + beta C, m, n, k ~ U{8..128}:
  a random sample of 300 products against the oracle, bit for bit in exact
  mode and within 4 k eps (|alpha| ||A|| ||B|| + |beta| ||C0||) in FAST mode.
* 50,000 SPD solves of order 32 and their adjoint: a random sample of 200
  matrices bit for bit against the oracle (factor, x, dL/dA, dL/db).
* 8,000 mixed-precision and plain FP64 LU solves, n in {128, 256, 384, 512}:
  every backward error recomputed in FP64, a sample against LAPACK, and the
  blocked FP64 factors of a sample by the dget01 ratio.
"""
import numpy as np
import pytest
import torch

import bench
import oracle
from gpu_util import assert_bits_equal

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def hipmod():
    from matflat_amd import hip
    return hip


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_spd_solve_100k_backward_error(mode):
    hip = hipmod()
    dev = torch.device("cuda")
    count = 100_000
    buf, offsets, ns, ldas, xoff, b = bench.make_spd_batch(count, 10, 100, seed=2)
    a0, b0 = buf.clone(), b.clone()
    n_t = torch.from_numpy(ns).to(dev)
    lda_t = torch.from_numpy(ldas).to(dev)
    a_ptrs = (buf.data_ptr() + 8 * torch.from_numpy(offsets).to(dev)).contiguous()
    x_ptrs = (b.data_ptr() + 8 * torch.from_numpy(xoff).to(dev)).contiguous()
    info = torch.full((count,), 99, dtype=torch.int32, device=dev)
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST if mode == "fast" else hip.MATH_EXACT)
    hip.dposv_vbatched(n_t, a_ptrs, lda_t, x_ptrs, info, 100, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert int((info != 0).sum().item()) == 0
    worst = 0.0
    for n in range(10, 101):
        idx = np.flatnonzero(ns == n)
        if idx.size == 0:
            continue
        c, lda = idx.size, int(ldas[idx[0]])
        start = int(offsets[idx].min())
        # matrices of one order are contiguous, in batch order: [matrix, column, row]
        orig = a0[start:start + c * n * lda].view(c, n, lda)
        fact = buf[start:start + c * n * lda].view(c, n, lda)
        at = orig[:, :, :n].transpose(1, 2)  # [matrix, row, column]
        upper = torch.triu(torch.ones(n, n, dtype=torch.bool, device=dev))
        a_full = torch.where(upper, at, at.transpose(1, 2))
        xi = torch.from_numpy(xoff[idx][:, None] + np.arange(n)[None, :]).to(dev)
        x, bb = b[xi], b0[xi]
        r = bb - torch.bmm(a_full, x.unsqueeze(2)).squeeze(2)
        eta = r.abs().amax(1) / (a_full.abs().sum(2).amax(1) * x.abs().amax(1) + bb.abs().amax(1))
        worst = max(worst, float((eta / (n * EPS)).max().item()))
        lt = fact[:, :, :n].transpose(1, 2)  # factor as [matrix, row, column]
        strict_upper = torch.triu(torch.ones(n, n, dtype=torch.bool, device=dev), 1)
        assert bool((lt[:, strict_upper] == 0).all()), n
        if lda > n:
            assert bool(torch.isnan(fact[:, :, n:]).all()), n  # padding rows never written
    assert worst <= 1.0, worst
    if mode == "exact":
        rng = np.random.default_rng(7)
        host_a0, host_b0 = a0.cpu().numpy(), b0.cpu().numpy()
        host_a, host_b = buf.cpu().numpy(), b.cpu().numpy()
        for i in rng.choice(count, 200, replace=False):
            n, lda, o, xo = int(ns[i]), int(ldas[i]), int(offsets[i]), int(xoff[i])
            wa = host_a0[o:o + lda * n].copy()
            wx = host_b0[xo:xo + n].copy()
            assert oracle.dposv(n, wa, lda, wx) == 0
            assert_bits_equal(host_a[o:o + lda * n], wa, f"factor {i} n={n}")
            assert_bits_equal(host_b[xo:xo + n], wx, f"x {i} n={n}")


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_gemm_200k_sample_against_oracle(mode):
    hip = hipmod()
    dev = torch.device("cuda")
    count = 200_000
    rng = np.random.default_rng(5)
    m, n, k = (rng.integers(8, 129, size=count).astype(np.int64) for _ in range(3))
    ta, tb = rng.choice([111, 112], size=count), rng.choice([111, 112], size=count)
    ra, ca = np.where(ta == 111, m, k), np.where(ta == 111, k, m)
    rb, cb = np.where(tb == 111, k, n), np.where(tb == 111, n, k)
    lda, ldb, ldc = (ra + 3) // 4 * 4, (rb + 3) // 4 * 4, (m + 3) // 4 * 4

    def slab(sizes):
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        t = torch.randn(int(sizes.sum()), dtype=torch.float64, device=dev)
        return t, off, (t.data_ptr() + 8 * torch.from_numpy(off).to(dev)).contiguous()

    A, aoff, pa = slab(lda * ca)
    B, boff, pb = slab(ldb * cb)
    C, coff, pc = slab(ldc * n)
    C0 = C.clone()
    alpha = torch.from_numpy(rng.uniform(-1, 1, count)).to(dev)
    beta = torch.from_numpy(rng.uniform(-1, 1, count)).to(dev)

    def i32(v):
        return torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).to(dev)

    info = torch.full((count,), 99, dtype=torch.int32, device=dev)
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST if mode == "fast" else hip.MATH_EXACT)
    hip.gemm_vbatched(i32(ta), i32(tb), i32(m), i32(n), i32(k), alpha, pa, i32(lda), pb, i32(ldb), beta, pc,
                      i32(ldc), info, 128, 128, 128, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert int((info != 0).sum().item()) == 0
    ha, hb, hc, hc0 = A.cpu().numpy(), B.cpu().numpy(), C.cpu().numpy(), C0.cpu().numpy()
    al, be = alpha.cpu().numpy(), beta.cpu().numpy()
    for i in np.random.default_rng(11).choice(count, 300, replace=False):
        a = ha[aoff[i]:aoff[i] + lda[i] * ca[i]]
        bm = hb[boff[i]:boff[i] + ldb[i] * cb[i]]
        want = hc0[coff[i]:coff[i] + ldc[i] * n[i]].copy()
        oracle.dgemm(int(ta[i]), int(tb[i]), int(m[i]), int(n[i]), int(k[i]), float(al[i]), a, int(lda[i]), bm,
                     int(ldb[i]), float(be[i]), want, int(ldc[i]))
        got = hc[coff[i]:coff[i] + ldc[i] * n[i]]
        if mode == "exact":
            assert_bits_equal(got, want, f"product {i}")
        else:
            def view(v, rows, cols, ld):
                return v.reshape(cols, ld)[:, :rows].T

            bound = 4 * k[i] * EPS * (abs(al[i]) * np.linalg.norm(view(a, ra[i], ca[i], lda[i]))
                                      * np.linalg.norm(view(bm, rb[i], cb[i], ldb[i]))
                                      + abs(be[i]) * np.linalg.norm(view(hc0[coff[i]:coff[i] + ldc[i] * n[i]],
                                                                         m[i], n[i], ldc[i])))
            err = np.linalg.norm(view(got, m[i], n[i], ldc[i]) - view(want, m[i], n[i], ldc[i]))
            assert err <= bound, (i, err, bound)
            # rows between m and ldc are never written
            pad = view(got, ldc[i], n[i], ldc[i])[m[i]:]
            assert np.array_equal(pad, view(hc0[coff[i]:coff[i] + ldc[i] * n[i]], ldc[i], n[i], ldc[i])[m[i]:])


def test_spd_solve_adjoint_50k_n32_sample_against_oracle():
    """The differentiated SPD-solve workload at full size: 50,000 matrices of
    order 32 (A = B B^T + 32 I), forward exact SPD solve, then the adjoint
    for loss = w . x.  Every info is 0, and a random sample of 200 matrices
    gives factor, x, dL/dA and dL/db bit-identical to the oracle."""
    hip = hipmod()
    dev = torch.device("cuda")
    count, n = 50_000, 32
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    bm = torch.randn((count, n, n), dtype=torch.float64, device=dev, generator=g)
    a = torch.bmm(bm, bm.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device=dev)
    del bm
    A = a.transpose(1, 2).contiguous().reshape(-1)  # column-major per matrix (symmetric anyway)
    A0 = A.clone()
    x = torch.randn(count * n, dtype=torch.float64, device=dev, generator=g)
    b0 = x.clone()
    w = torch.randn(count * n, dtype=torch.float64, device=dev, generator=g)
    abar = torch.zeros(count * n * n, dtype=torch.float64, device=dev)
    bbar = torch.zeros(count * n, dtype=torch.float64, device=dev)
    idx = torch.arange(count, dtype=torch.int64, device=dev)
    mat = lambda t: (t.data_ptr() + 8 * n * n * idx).contiguous()  # noqa: E731
    vec = lambda t: (t.data_ptr() + 8 * n * idx).contiguous()  # noqa: E731
    n_t = torch.full((count,), n, dtype=torch.int32, device=dev)
    info = torch.full((count,), 99, dtype=torch.int32, device=dev)
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_EXACT)
    hip.dposv_vbatched(n_t, mat(A), n_t, vec(x), info, n, handle=h)
    info2 = torch.full((count,), 99, dtype=torch.int32, device=dev)
    hip.dposv_backward_vbatched(n_t, mat(A), n_t, vec(x), vec(w), mat(abar), n_t, vec(bbar), info2, n, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert int((info != 0).sum().item()) == 0 and int((info2 != 0).sum().item()) == 0
    ha0, hb0, hw = A0.cpu().numpy(), b0.cpu().numpy(), w.cpu().numpy()
    ha, hx, hab, hbb = A.cpu().numpy(), x.cpu().numpy(), abar.cpu().numpy(), bbar.cpu().numpy()
    nn = n * n
    for i in np.random.default_rng(13).choice(count, 200, replace=False):
        la = ha0[i * nn:(i + 1) * nn].copy()
        lx = hb0[i * n:(i + 1) * n].copy()
        assert oracle.dposv(n, la, n, lx) == 0
        assert_bits_equal(ha[i * nn:(i + 1) * nn], la, f"factor {i}")
        assert_bits_equal(hx[i * n:(i + 1) * n], lx, f"x {i}")
        want_abar, want_bbar = oracle.dposv_backward(n, la, n, lx, hw[i * n:(i + 1) * n])
        assert_bits_equal(hab[i * nn:(i + 1) * nn], want_abar, f"abar {i}")
        assert_bits_equal(hbb[i * n:(i + 1) * n], want_bbar, f"bbar {i}")


def _group_residuals(buf, offsets, ns, b, boff, x):
    """eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) of every
    solve, recomputed on the GPU in FP64 from the original A (column-major,
    lda = n), per order."""
    dev = torch.device("cuda")
    eta = np.zeros(ns.size)
    for n in sorted(set(ns.tolist())):
        idx = np.flatnonzero(ns == n)
        ao = torch.from_numpy(offsets[idx]).to(dev)
        bo = torch.from_numpy(boff[idx]).to(dev)
        ar = torch.arange(n * n, device=dev)
        vr = torch.arange(n, device=dev)
        for s0 in range(0, idx.size, 500):
            sl = slice(s0, s0 + 500)
            a = buf[ao[sl, None] + ar[None, :]].view(-1, n, n).transpose(1, 2).contiguous()  # row-major A
            xv = x[bo[sl, None] + vr[None, :]]
            bv = b[bo[sl, None] + vr[None, :]]
            r = bv - torch.bmm(a, xv.unsqueeze(2)).squeeze(2)
            den = a.abs().sum(dim=2).amax(dim=1) * xv.abs().amax(dim=1) + bv.abs().amax(dim=1)
            eta[idx[sl]] = (r.abs().amax(dim=1) / den).cpu().numpy()
    return eta


def test_mixed_precision_8000_solves_backward_error():
    """The mixed-precision workload at full size (8,000 LU solves, n uniform
    over {128, 256, 384, 512}, cond 10^U(2,6), built by
    bench.make_conditioned_batch with seed 3), both solvers: every solve has
    info = 0 and a backward error, recomputed here in FP64 from the original
    A, <= 1e-15 n; the mixed path took <= 10 refinement steps; each device
    berr agrees with the recomputation; a sample of 50 solutions matches
    LAPACK (numpy.linalg.solve) within cond * n * 20 eps; a sample of 24
    matrices factored by the blocked FP64 LU meets
    ||PA - LU||_1 / (||A||_1 n eps) <= 10."""
    hip = hipmod()
    dev = torch.device("cuda")
    count, seed = 8000, 3
    buf, offsets, ns, b, boff = bench.make_conditioned_batch(count, seed)
    rng = np.random.default_rng(seed)
    rng.choice(np.array(bench.MIXED_SIZES), size=count)
    kappa = 10.0 ** rng.uniform(2.0, 6.0, size=count)
    n_t = torch.from_numpy(ns).to(dev)
    a_ptrs = (buf.data_ptr() + 8 * torch.from_numpy(offsets).to(dev)).contiguous()
    b_ptrs = (b.data_ptr() + 8 * torch.from_numpy(boff).to(dev)).contiguous()
    a0 = buf.clone()
    sample = np.random.default_rng(17).choice(count, size=50, replace=False)
    for mixed in (True, False):
        x = torch.zeros_like(b)
        x_ptrs = (x.data_ptr() + 8 * torch.from_numpy(boff).to(dev)).contiguous()
        info = torch.full((count,), 99, dtype=torch.int32, device=dev)
        iters = torch.full((count,), -1, dtype=torch.int32, device=dev)
        berr = torch.full((count,), -1.0, dtype=torch.float64, device=dev)
        if mixed:
            hip.dgesv_mixed_vbatched(n_t, a_ptrs, n_t, b_ptrs, x_ptrs, 10, 1e-15, iters, berr, info, 512)
        else:
            hip.dgesv_vbatched(n_t, a_ptrs, n_t, b_ptrs, x_ptrs, berr, info, 512)
        torch.cuda.synchronize()
        assert bool(torch.equal(buf, a0)), "A must not be modified"
        inf = info.cpu().numpy()
        assert np.all(inf == 0), (mixed, np.flatnonzero(inf)[:10])
        eta = _group_residuals(buf, offsets, ns, b, boff, x)
        assert np.all(eta <= 1e-15 * ns), (mixed, float((eta / ns).max()))
        be = berr.cpu().numpy()
        assert np.all(np.abs(be - eta) <= 0.5 * eta + 4 * EPS), (mixed, float(np.abs(be - eta).max()))
        if mixed:
            it = iters.cpu().numpy()
            assert np.all((it >= 0) & (it <= 10)), (int(it.min()), int(it.max()))
        xh, bh, ah = x.cpu().numpy(), b.cpu().numpy(), buf.cpu().numpy()
        for i in sample:
            n = int(ns[i])
            a = ah[offsets[i]:offsets[i] + n * n].reshape(n, n).T
            want = np.linalg.solve(a, bh[boff[i]:boff[i] + n])
            got = xh[boff[i]:boff[i] + n]
            fwd = np.abs(got - want).max() / np.abs(want).max()
            assert fwd <= kappa[i] * n * 20 * EPS, (mixed, i, n, kappa[i], fwd)
    # the blocked FP64 factors themselves, on copies of 24 of the matrices
    fsample = np.random.default_rng(23).choice(count, size=24, replace=False)
    ah = a0.cpu().numpy()
    sub = [ah[offsets[i]:offsets[i] + int(ns[i]) ** 2].copy() for i in fsample]
    from gpu_util import Packed, i32
    A = Packed(sub)
    P = Packed([np.full(int(ns[i]), -1, dtype=np.int32) for i in fsample], dtype=torch.int32)
    sn = [int(ns[i]) for i in fsample]
    finfo = i32(np.full(len(sn), 99))
    hip.dgetrf_blocked_vbatched(i32(sn), A.ptrs, i32(sn), P.ptrs, finfo, 512)
    torch.cuda.synchronize()
    assert np.all(finfo.cpu().numpy() == 0)
    for k, (lu_buf, piv) in enumerate(zip(A.all(), P.all())):
        n = sn[k]
        a = sub[k].reshape(n, n).T
        lu = lu_buf.reshape(n, n).T
        low, up = np.tril(lu, -1) + np.eye(n), np.triu(lu)
        assert sorted(piv.tolist()) == list(range(n))
        ratio = np.abs(a[piv, :] - low @ up).sum(axis=0).max() / (np.abs(a).sum(axis=0).max() * n * EPS)
        assert ratio <= 10.0, (k, n, ratio)
