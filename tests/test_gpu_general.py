"""GPU parity of the generic MatFlat-order kernels (mfh_general.hip):
complex128 LU, Cholesky and triangular solve at every order, and the double
LU, Cholesky, SPD solve, triangular solve and LU solve above the LDS-resident
range (129 .. MFH_EXACT_MAX_N rows).  Every output is compared bit for bit
with the oracle (oracle/matflat_oracle.c), which follows
Factorization.Lu.cs:296-406, Factorization.Cholesky.cs:114-249 and
Blas.SolveTriangular.cs:162-445."""
import os

import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, i32
from helpers import hpd_upper_cbuf, rand_buf, rand_cbuf, spd_upper_buf

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hipmod():
    from matflat_amd import hip
    return hip


def exact_handle():
    hip = hipmod()
    h = hip.Handle()
    h.set_math_mode(hip.MATH_EXACT)
    h.bind_stream()
    return h


# ------------------------------------------------------------- complex LU
def check_zgetrf(cases, bufs, what, max_m=None, max_n=None):
    hip = hipmod()
    print(f"[zgetrf] {what}: {len(cases)} matrices", flush=True)
    A = Packed(bufs, dtype=torch.complex128)
    P = Packed([np.full(m, -7, dtype=np.int32) for m, _, _ in cases], dtype=torch.int32)
    sign = i32(np.zeros(len(cases)))
    info = i32(np.full(len(cases), 99))
    ms, ns, ldas = [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases]
    h = exact_handle()
    hip.zgetrf_vbatched(i32(ms), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, max_m or max(ms),
                        max_n or max(ns), handle=h)
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    ga, gp, gs = A.all(), P.all(), sign.cpu().numpy()
    for i, (m, n, lda) in enumerate(cases):
        want = bufs[i].copy()
        wpiv, wsign = oracle.zgetrf(m, n, want, lda)
        assert_bits_equal(ga[i], want, f"{what} zgetrf {i} {(m, n, lda)}")
        assert np.array_equal(gp[i], wpiv), (what, i)
        assert gs[i] == wsign, (what, i)


def test_golden_complex_lu():
    g = np.load(os.path.join(GOLDEN, "complex.npz"), allow_pickle=False)
    hip = hipmod()
    meta = g["zlu_meta"]
    ms, ns, ldas = meta[:, 0], meta[:, 1], meta[:, 2]
    bufs = [g["zlu_in"][o:o + l * n].copy() for o, l, n in zip(g["zlu_off"], ldas, ns)]
    A = Packed(bufs, dtype=torch.complex128)
    P = Packed([np.full(m, -7, dtype=np.int32) for m in ms], dtype=torch.int32)
    sign = i32(np.zeros(len(ms)))
    info = i32(np.full(len(ms), 99))
    hip.zgetrf_vbatched(i32(ms), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, int(ms.max()), int(ns.max()),
                        handle=exact_handle())
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    ga, gp, gs = A.all(), P.all(), sign.cpu().numpy()
    for i, (m, n, lda, sg) in enumerate(meta):
        o, po = g["zlu_off"][i], g["zlu_piv_off"][i]
        assert_bits_equal(ga[i], g["zlu_out"][o:o + lda * n], f"zlu {i}")
        assert np.array_equal(gp[i], g["zlu_piv"][po:po + m]) and gs[i] == sg, i


def test_golden_complex_cholesky_and_trsv():
    g = np.load(os.path.join(GOLDEN, "complex.npz"), allow_pickle=False)
    hip = hipmod()
    meta = g["zchol_meta"]
    ns, ldas = meta[:, 0], meta[:, 1]
    bufs = [g["zchol_in"][o:o + l * n].copy() for o, l, n in zip(g["zchol_off"], ldas, ns)]
    A = Packed(bufs, dtype=torch.complex128)
    info = i32(np.full(len(ns), 99))
    hip.zpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, int(ns.max()), handle=exact_handle())
    torch.cuda.synchronize()
    assert np.array_equal(info.cpu().numpy(), meta[:, 2])
    for i, got in enumerate(A.all()):
        o = g["zchol_off"][i]
        assert_bits_equal(got, g["zchol_out"][o:o + ldas[i] * ns[i]], f"zchol {i}")
    tm = g["ztrsv_meta"]
    for form in sorted({(int(u), int(t)) for u, t in tm[:, :2]}):
        rows = [i for i in range(len(tm)) if (tm[i, 0], tm[i, 1]) == form]
        a = [g["ztrsv_a"][g["ztrsv_off"][i]:g["ztrsv_off"][i] + tm[i, 3] * tm[i, 2]].copy() for i in rows]
        x = [g["ztrsv_x_in"][g["ztrsv_x_off"][i]:g["ztrsv_x_off"][i] + tm[i, 2] * tm[i, 4]].copy() for i in rows]
        A = Packed(a, dtype=torch.complex128)
        X = Packed(x, dtype=torch.complex128)
        info = i32(np.full(len(rows), 99))
        hip.ztrsv_vbatched(form[0], form[1], i32(tm[rows, 2]), A.ptrs, i32(tm[rows, 3]), X.ptrs, i32(tm[rows, 4]),
                           info, int(tm[rows, 2].max()), handle=exact_handle())
        torch.cuda.synchronize()
        assert np.all(info.cpu().numpy() == 0)
        for k, i in enumerate(rows):
            xo = g["ztrsv_x_off"][i]
            assert_bits_equal(X.get(k), g["ztrsv_x_out"][xo:xo + tm[i, 2] * tm[i, 4]], f"ztrsv {form} {i}")


@pytest.mark.parametrize("top", [40, 70, 300])
def test_zlu_variable_sizes(top):
    """top 40: staged in LDS; 70 and 300: worked on in place in HBM."""
    rng = np.random.default_rng(top)
    cases = []
    for _ in range(60 if top < 300 else 12):
        m, n = int(rng.integers(1, top + 1)), int(rng.integers(1, top + 1))
        cases.append((m, n, m + int(rng.integers(0, 4))))
    bufs = [rand_cbuf(top * 100 + i, m, n, lda, dist="normal") for i, (m, n, lda) in enumerate(cases)]
    check_zgetrf(cases, bufs, f"var{top}")


def test_zlu_special_matrices():
    nan, inf = np.nan, np.inf
    cols = [np.zeros(9, dtype=np.complex128),
            np.arange(1.0, 10.0) * (1 + 1j),
            np.array([1 + 2j, 3, -2 - 1j, 0, 1, 0, 0, 0, 1]),  # tie in |re| + |im|
            np.array([1, complex(nan, 0), 2j, 0, 1, 0, 0, 0, 1]),
            np.array([complex(0, nan), 5, 7, 0, 1, 0, 0, 0, 1]),
            np.array([1, complex(inf, 1), 2, 1, 2, 3, 4, 5, 7])]
    check_zgetrf([(3, 3, 3)] * len(cols), [c.astype(np.complex128) for c in cols], "special")


# ------------------------------------------------------- complex Cholesky
@pytest.mark.parametrize("top", [30, 150])
def test_zcholesky_variable_sizes_and_failures(top):
    """top 30: L staged in LDS; 150: in the per-workgroup HBM scratch.
    Every fourth matrix has a negative diagonal entry: info = j + 1 and the
    partial writes must match."""
    hip = hipmod()
    rng = np.random.default_rng(top + 1)
    ns = [int(v) for v in rng.integers(1, top + 1, size=50 if top < 100 else 16)]
    ns[0] = top
    ldas = [n + int(rng.integers(0, 5)) for n in ns]
    bufs = []
    for i, (n, lda) in enumerate(zip(ns, ldas)):
        b = hpd_upper_cbuf(top * 1000 + i, n, lda)
        if i % 4 == 3 and n > 1:
            f = int(rng.integers(0, n))
            b[f * lda + f] = -1.0 - 1j
        bufs.append(b)
    print(f"[zpotrf] top {top}: {len(ns)} matrices", flush=True)
    A = Packed(bufs, dtype=torch.complex128)
    info = i32(np.full(len(ns), 99))
    hip.zpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=exact_handle())
    torch.cuda.synchronize()
    gi = info.cpu().numpy()
    for i, got in enumerate(A.all()):
        want = bufs[i].copy()
        winfo = oracle.zpotrf(ns[i], want, ldas[i])
        assert gi[i] == winfo, (i, ns[i])
        assert_bits_equal(got, want, f"zpotrf {i} n={ns[i]}")
    assert np.any(gi > 0) and np.any(gi == 0)


# ------------------------------------------------- complex triangular solve
@pytest.mark.parametrize("top", [50, 200])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112, 113, 114])
def test_ztrsv_all_forms(top, uplo, trans):
    hip = hipmod()
    rng = np.random.default_rng(top + uplo + trans)
    ns = [int(v) for v in rng.integers(1, top + 1, size=24)]
    ns[0] = top
    ldas = [n + int(rng.integers(0, 3)) for n in ns]
    incs = [int(rng.integers(1, 4)) for _ in ns]
    abufs, xbufs = [], []
    for i, (n, lda) in enumerate(zip(ns, ldas)):
        a = rand_cbuf(i + 7 * top, lda, n, lda)
        for j in range(n):
            a[j * lda + j] = (2.0 + np.sqrt(n)) * np.exp(1j * rng.uniform(0, 2 * np.pi))
        abufs.append(a)
        xbufs.append(rand_cbuf(i + 11 * top, n * incs[i], 1, n * incs[i]))
    A = Packed(abufs, dtype=torch.complex128)
    X = Packed(xbufs, dtype=torch.complex128)
    info = i32(np.full(len(ns), 99))
    hip.ztrsv_vbatched(uplo, trans, i32(ns), A.ptrs, i32(ldas), X.ptrs, i32(incs), info, max(ns),
                       handle=exact_handle())
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    for i, got in enumerate(X.all()):
        want = xbufs[i].copy()
        oracle.ztrsv(uplo, trans, ns[i], abufs[i], ldas[i], want, incs[i])
        assert_bits_equal(got, want, f"ztrsv {uplo} {trans} {i} n={ns[i]}")


def test_complex_invalid_arguments():
    hip = hipmod()
    a = Packed([np.ones(16, dtype=np.complex128)] * 3, dtype=torch.complex128)
    info = i32(np.full(3, 99))
    piv = Packed([np.zeros(4, dtype=np.int32)] * 3, dtype=torch.int32)
    sign = i32(np.zeros(3))
    # m = 0; n beyond max_n; lda < m
    hip.zgetrf_vbatched(i32([0, 4, 4]), i32([4, 9, 4]), a.ptrs, i32([4, 4, 3]), piv.ptrs, sign, info, 4, 4,
                        handle=exact_handle())
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [-3, -4, -6]
    info = i32(np.full(3, 99))
    hip.zpotrf_vbatched(i32([0, 4, 4]), a.ptrs, i32([4, 4, 2]), info, 4, handle=exact_handle())
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [-3, 2, -5]  # the all-ones matrix fails at row 1
    info = i32(np.full(3, 99))
    hip.ztrsv_vbatched(121, 113, i32([4, 4, 4]), a.ptrs, i32([4, 3, 4]), a.ptrs, i32([1, 1, 0]), info, 4,
                       handle=exact_handle())
    torch.cuda.synchronize()
    assert info.cpu().tolist()[1:] == [-7, -9]
    with pytest.raises(hip.ArgumentError):
        hip.ztrsv_vbatched(121, 115, i32([4]), a.ptrs, i32([4]), a.ptrs, i32([1]), info, 4)


# ------------------------------------------------ double above 128 rows
def test_dgetrf_large_orders_mixed_with_small():
    """Sizes on both sides of the LDS kernels' range in one call: the small
    ones go through the size sort, the large ones through k_getrf_gen."""
    hip = hipmod()
    cases = [(129, 129, 129), (200, 150, 203), (150, 300, 151), (256, 256, 256), (511, 511, 512), (512, 512, 512),
             (700, 90, 700), (64, 64, 64), (5, 3, 7), (128, 128, 130), (1024, 1024, 1024)]
    bufs = [rand_buf(40 + i, m, n, lda, dist="normal") for i, (m, n, lda) in enumerate(cases)]
    print(f"[dgetrf large] {len(cases)} matrices", flush=True)
    A = Packed(bufs)
    P = Packed([np.full(m, -7, dtype=np.int32) for m, _, _ in cases], dtype=torch.int32)
    sign = i32(np.zeros(len(cases)))
    info = i32(np.full(len(cases), 99))
    ms, ns, ldas = [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases]
    hip.dgetrf_vbatched(i32(ms), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, max(ms), max(ns),
                        handle=exact_handle())
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    ga, gp, gs = A.all(), P.all(), sign.cpu().numpy()
    for i, (m, n, lda) in enumerate(cases):
        want = bufs[i].copy()
        wpiv, wsign = oracle.dgetrf(m, n, want, lda)
        assert_bits_equal(ga[i], want, f"dgetrf {i} {(m, n, lda)}")
        assert np.array_equal(gp[i], wpiv) and gs[i] == wsign, i


def test_dgetrf_strided_large():
    hip = hipmod()
    count, n = 6, 384
    a = np.random.default_rng(3).standard_normal(count * n * n)
    dev = torch.from_numpy(a.copy()).cuda()
    piv = torch.zeros(count * n, dtype=torch.int32, device="cuda")
    sign = torch.zeros(count, dtype=torch.int32, device="cuda")
    info = torch.full((count,), 99, dtype=torch.int32, device="cuda")
    hip.dgetrf_strided_batched(count, n, n, dev, n, n * n, piv, n, sign, info, handle=exact_handle())
    torch.cuda.synchronize()
    wpiv, wsign = oracle.dgetrf_batch(count, n, n, n, n * n, a, threads=8)
    assert_bits_equal(dev.cpu().numpy(), a, "strided n=384")
    assert np.array_equal(piv.cpu().numpy(), wpiv) and np.array_equal(sign.cpu().numpy(), wsign)
    assert torch.all(info == 0)


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_dpotrf_dposv_large_orders(math):
    """Cholesky and the SPD solve at 129..520 rows beside small matrices;
    FAST mode runs the MatFlat-order kernels above its register tiles, so
    the large ones are bit-identical there too.  Two matrices fail."""
    hip = hipmod()
    h = exact_handle()
    if math == "fast":
        h.set_math_mode(hip.MATH_FAST)
    ns = [129, 200, 256, 300, 511, 512, 520, 17, 100, 300, 150]
    ldas = [(n + 7) // 8 * 8 for n in ns]
    bufs = [spd_upper_buf(300 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    bufs[9][200 * ldas[9] + 200] = -5.0  # fails at row 200
    bufs[10][ldas[10] * 0] = np.nan      # NaN diagonal: fails at row 0
    xs = [np.random.default_rng(9 + i).standard_normal(n) for i, n in enumerate(ns)]
    print(f"[dpotrf/dposv large] {math}", flush=True)
    A = Packed(bufs)
    info = i32(np.full(len(ns), 99))
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=h)
    torch.cuda.synchronize()
    gi = info.cpu().numpy()
    large = [i for i, n in enumerate(ns) if n > 128]
    for i, got in enumerate(A.all()):
        want = bufs[i].copy()
        winfo = oracle.dpotrf(ns[i], want, ldas[i])
        assert gi[i] == winfo, (i, ns[i])
        if math == "exact" or i in large:
            assert_bits_equal(got, want, f"dpotrf {i} n={ns[i]}")
    A2 = Packed(bufs)
    X = Packed(xs)
    info = i32(np.full(len(ns), 99))
    hip.dposv_vbatched(i32(ns), A2.ptrs, i32(ldas), X.ptrs, info, max(ns), handle=h)
    torch.cuda.synchronize()
    gi = info.cpu().numpy()
    ga, gx = A2.all(), X.all()
    for i, n in enumerate(ns):
        wa, wx = bufs[i].copy(), xs[i].copy()
        winfo = oracle.dposv(n, wa, ldas[i], wx)
        assert gi[i] == winfo, (i, n)
        if math == "exact" or i in large:
            assert_bits_equal(ga[i], wa, f"dposv factor {i}")
            assert_bits_equal(gx[i], wx, f"dposv x {i}")


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("diag", [131, 132])
def test_dtrsv_large_orders(uplo, trans, diag):
    hip = hipmod()
    rng = np.random.default_rng(uplo + trans + diag)
    ns = [129, 257, 400, 512, 1000, 33, 128]
    ldas = [n + 3 for n in ns]
    incs = [1, 2, 3, 1, 1, 2, 1]
    abufs, xbufs = [], []
    for i, (n, lda) in enumerate(zip(ns, ldas)):
        a = rng.uniform(-1, 1, lda * n)
        for j in range(n):
            a[j * lda + j] = 2.0 + np.sqrt(n) * abs(a[j * lda + j])
        abufs.append(a)
        xbufs.append(rng.uniform(-1, 1, n * incs[i]))
    A = Packed(abufs)
    X = Packed(xbufs)
    info = i32(np.full(len(ns), 99))
    hip.dtrsv_vbatched(uplo, trans, diag, i32(ns), A.ptrs, i32(ldas), X.ptrs, i32(incs), info, max(ns),
                       handle=exact_handle())
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    for i, got in enumerate(X.all()):
        a = abufs[i].copy()
        if diag == 132:
            for j in range(ns[i]):
                a[j * ldas[i] + j] = 1.0
        want = xbufs[i].copy()
        oracle.dtrsv(uplo, trans, ns[i], a, ldas[i], want, incs[i])
        assert_bits_equal(got, want, f"dtrsv {uplo} {trans} {diag} n={ns[i]}")


def test_dgetrs_large_orders_after_getrf():
    hip = hipmod()
    ns = [129, 300, 512, 64, 777]
    ldas = [n + 1 for n in ns]
    bufs = [rand_buf(70 + i, n, n, l, dist="normal") for i, (n, l) in enumerate(zip(ns, ldas))]
    bs = [np.random.default_rng(80 + i).standard_normal(n) for i, n in enumerate(ns)]
    h = exact_handle()
    A = Packed(bufs)
    P = Packed([np.zeros(n, dtype=np.int32) for n in ns], dtype=torch.int32)
    sign = i32(np.zeros(len(ns)))
    info = i32(np.full(len(ns), 99))
    hip.dgetrf_vbatched(i32(ns), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, max(ns), max(ns), handle=h)
    Bv = Packed(bs)
    Xv = Packed([np.zeros(n) for n in ns])
    info2 = i32(np.full(len(ns), 99))
    hip.dgetrs_vbatched(i32(ns), A.ptrs, i32(ldas), P.ptrs, Bv.ptrs, Xv.ptrs, info2, max(ns), handle=h)
    # b and x aliased
    info3 = i32(np.full(len(ns), 99))
    hip.dgetrs_vbatched(i32(ns), A.ptrs, i32(ldas), P.ptrs, Bv.ptrs, Bv.ptrs, info3, max(ns), handle=h)
    torch.cuda.synchronize()
    assert np.all(info2.cpu().numpy() == 0) and np.all(info3.cpu().numpy() == 0)
    gx, gb = Xv.all(), Bv.all()
    for i, n in enumerate(ns):
        lu = bufs[i].copy()
        piv, _ = oracle.dgetrf(n, n, lu, ldas[i])
        want = oracle.dgetrs(n, lu, ldas[i], piv, bs[i])
        assert_bits_equal(gx[i], want, f"dgetrs {i} n={n}")
        assert_bits_equal(gb[i], want, f"dgetrs aliased {i} n={n}")


def test_limits_of_small_entry_points():
    hip = hipmod()
    h = exact_handle()
    a = Packed([np.ones(4)])
    info = i32([99])
    with pytest.raises(hip.ArgumentError):
        hip.dposv_vbatched_noalloc(i32([2]), a.ptrs, i32([2]), a.ptrs, info, 129, handle=h)
    with pytest.raises(hip.ArgumentError):
        hip.dgetri_vbatched(i32([2]), a.ptrs, i32([2]), a.ptrs, info, 129, handle=h)


def test_matflat_complex_overloads():
    """The drop-in signatures with complex128 tensors select the Complex
    overloads (Factorization.Lu.cs:199-238, Factorization.Cholesky.cs:195-249,
    Blas.SolveTriangular.cs:280-395)."""
    from matflat_amd.matflat import (ArgumentException, Blas, Factorization, MatrixFactorizationException,
                                     Transpose, Uplo)
    m, n, lda = 13, 9, 15
    a = rand_cbuf(5, m, n, lda, dist="normal")
    dev = torch.from_numpy(a.copy()).cuda()
    piv = torch.zeros(m, dtype=torch.int32, device="cuda")
    sign = Factorization.Lu(m, n, dev, lda, piv)
    want = a.copy()
    wpiv, wsign = oracle.zgetrf(m, n, want, lda)
    assert_bits_equal(dev.cpu().numpy(), want, "complex Lu")
    assert np.array_equal(piv.cpu().numpy(), wpiv) and sign == wsign
    h = hpd_upper_cbuf(8, 7, 8)
    dev = torch.from_numpy(h.copy()).cuda()
    Factorization.Cholesky(7, dev, 8)
    want = h.copy()
    assert oracle.zpotrf(7, want, 8) == 0
    assert_bits_equal(dev.cpu().numpy(), want, "complex Cholesky")
    bad = h.copy()
    bad[3 * 8 + 3] = -1.0
    dev = torch.from_numpy(bad.copy()).cuda()
    with pytest.raises(MatrixFactorizationException):
        Factorization.Cholesky(7, dev, 8)
    want = bad.copy()
    assert oracle.zpotrf(7, want, 8) == 4
    assert_bits_equal(dev.cpu().numpy(), want, "complex Cholesky partial writes")
    t = rand_cbuf(9, 6, 6, 6) + 3.0 * np.eye(6).reshape(-1)
    tri = torch.from_numpy(t.copy()).cuda()
    for tr in (Transpose.NoTrans, Transpose.Trans, Transpose.ConjTrans, Transpose.ConjNoTrans):
        x0 = rand_cbuf(10 + int(tr), 6, 1, 6)
        x = torch.from_numpy(x0.copy()).cuda()
        Blas.SolveTriangular(Uplo.Upper, tr, 6, tri, 6, x, 1)
        want = x0.copy()
        oracle.ztrsv(121, int(tr), 6, t, 6, want, 1)
        assert_bits_equal(x.cpu().numpy(), want, f"complex SolveTriangular {tr}")
    with pytest.raises(ArgumentException):
        Blas.SolveTriangular(Uplo.Upper, 115, 6, tri, 6, x, 1)
    with pytest.raises(ArgumentException):
        Blas.SolveTriangular(Uplo.Lower, Transpose.ConjTrans, 6, torch.zeros(36, dtype=torch.float64,
                                                                                  device="cuda"), 6,
                             torch.zeros(6, dtype=torch.float64, device="cuda"), 1)
