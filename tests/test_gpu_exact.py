"""GPU parity of the exact kernels against the CPU oracle (MatFlat's operation
order restated in oracle/matflat_oracle.c).  The bar is bit-identical output:
factors, pivots, signs, solutions and untouched padding."""
import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, f64, i32
from helpers import rand_buf, rand_cbuf, spd_upper_buf, view
from test_oracle import GEMM_CASES, LU_CASES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from matflat_amd import hip  # noqa: F401  (fails loudly if the library is missing)
    yield


def hipmod():
    from matflat_amd import hip
    return hip


# ------------------------------------------------------------------ LU
def run_getrf(cases, bufs):
    hip = hipmod()
    ms = [c[0] for c in cases]
    ns = [c[1] for c in cases]
    ldas = [c[2] for c in cases]
    A = Packed(bufs)
    P = Packed([np.full(m, -7, dtype=np.int32) for m in ms], dtype=torch.int32)
    sign = i32(np.zeros(len(cases)))
    info = i32(np.full(len(cases), 99))
    hip.dgetrf_vbatched(i32(ms), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, max(ms), max(ns))
    torch.cuda.synchronize()
    return A.all(), P.all(), sign.cpu().numpy(), info.cpu().numpy()


def check_getrf(cases, bufs, what):
    got_a, got_p, got_s, got_i = run_getrf(cases, [b.copy() for b in bufs])
    assert np.all(got_i == 0), got_i
    for idx, ((m, n, lda), b) in enumerate(zip(cases, bufs)):
        want = b.copy()
        piv, sign = oracle.dgetrf(m, n, want, lda)
        assert_bits_equal(got_a[idx], want, f"{what} A case {(m, n, lda)}")
        assert np.array_equal(got_p[idx], piv), (m, n, lda)
        assert got_s[idx] == sign, (m, n, lda)


def test_lu_reference_shapes_one_batch():
    bufs = [rand_buf(42, m, n, lda) for (m, n, lda) in LU_CASES]
    check_getrf(LU_CASES, bufs, "LuTests shapes")


def test_lu_special_matrices():
    cases = [(3, 3, 3)] * 6
    bufs = [np.zeros(9), np.arange(1.0, 10.0),
            np.array([np.nan, 5.0, 3.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0]),
            np.array([1.0, np.nan, 3.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0]),
            np.array([1.0, np.inf, -np.inf, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0]),
            np.array([1.0, -2.0, 2.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0])]
    check_getrf(cases, bufs, "special")


@pytest.mark.parametrize("seed", [0, 1])
def test_lu_variable_sizes(seed):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(300):
        m = int(rng.integers(1, 129))
        n = int(rng.integers(1, 129))
        lda = m + int(rng.integers(0, 5))
        cases.append((m, n, lda))
    bufs = [rand_buf(1000 * seed + i, m, n, lda, dist="normal") for i, (m, n, lda) in enumerate(cases)]
    check_getrf(cases, bufs, "var")


def test_lu_strided_n64_batch():
    hip = hipmod()
    count, n = 1024, 64
    rng = np.random.default_rng(1)
    a = rng.standard_normal(count * n * n)
    dev = torch.from_numpy(a.copy()).cuda()
    piv = torch.zeros(count * n, dtype=torch.int32, device="cuda")
    sign = torch.zeros(count, dtype=torch.int32, device="cuda")
    info = torch.full((count,), 99, dtype=torch.int32, device="cuda")
    hip.dgetrf_strided_batched(count, n, n, dev, n, n * n, piv, n, sign, info)
    torch.cuda.synchronize()
    want_piv, want_sign = oracle.dgetrf_batch(count, n, n, n, n * n, a, threads=8)
    assert_bits_equal(dev.cpu().numpy(), a, "n=64 batch")
    assert np.array_equal(piv.cpu().numpy(), want_piv)
    assert np.array_equal(sign.cpu().numpy(), want_sign)
    assert torch.all(info == 0)


# ------------------------------------------------------------------ Cholesky / posv
def chol_cases(seed, count, nmax=128, pad=8):
    rng = np.random.default_rng(seed)
    ns = rng.integers(1, nmax + 1, size=count)
    ldas = [int(n + rng.integers(0, pad)) for n in ns]
    return [int(n) for n in ns], ldas


def test_cholesky_variable_sizes_nan_lower():
    hip = hipmod()
    ns, ldas = chol_cases(3, 200)
    bufs = [spd_upper_buf(i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    A = Packed(bufs)
    info = i32(np.full(len(ns), 99))
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns))
    torch.cuda.synchronize()
    got = A.all()
    assert np.all(info.cpu().numpy() == 0)
    for i, (n, l) in enumerate(zip(ns, ldas)):
        want = bufs[i].copy()
        assert oracle.dpotrf(n, want, l) == 0
        assert_bits_equal(got[i], want, f"potrf n={n} lda={l}")


def test_cholesky_failures_match_partial_writes():
    hip = hipmod()
    bufs, ns = [], []
    for seed in range(20):
        n = 2 + seed
        rng = np.random.default_rng(seed)
        up = np.triu(2 * rng.random((n, n)) - 1) + np.diag(rng.random(n) * (seed % 3))
        buf = np.full(n * n, np.nan)
        for j in range(n):
            buf[j * n:j * n + j + 1] = up[: j + 1, j]
        bufs.append(buf)
        ns.append(n)
    bufs.append(np.array([np.nan, np.nan, 1.0, 1.0]))  # NaN diagonal
    ns.append(2)
    A = Packed(bufs)
    info = i32(np.full(len(ns), 99))
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ns), info, max(ns))
    torch.cuda.synchronize()
    got = A.all()
    gi = info.cpu().numpy()
    for i, n in enumerate(ns):
        want = bufs[i].copy()
        winfo = oracle.dpotrf(n, want, n)
        assert gi[i] == winfo, (i, n)
        assert_bits_equal(got[i], want, f"potrf failure case {i}")


def test_posv_variable_sizes():
    hip = hipmod()
    ns, ldas = chol_cases(5, 300, nmax=128)
    bufs = [spd_upper_buf(100 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    xs = [np.random.default_rng(7 + i).standard_normal(n) for i, n in enumerate(ns)]
    A = Packed(bufs)
    X = Packed(xs)
    info = i32(np.full(len(ns), 99))
    hip.dposv_vbatched(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, max(ns))
    torch.cuda.synchronize()
    ga, gx = A.all(), X.all()
    assert np.all(info.cpu().numpy() == 0)
    for i, (n, l) in enumerate(zip(ns, ldas)):
        wa, wx = bufs[i].copy(), xs[i].copy()
        assert oracle.dposv(n, wa, l, wx) == 0
        assert_bits_equal(ga[i], wa, f"posv A n={n}")
        assert_bits_equal(gx[i], wx, f"posv x n={n}")


def test_posv_plan_replays_match_eager():
    hip = hipmod()
    ns, ldas = chol_cases(11, 500, nmax=100)
    bufs = [spd_upper_buf(300 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    xs = [np.random.default_rng(70 + i).standard_normal(n) for i, n in enumerate(ns)]
    A0, X0 = Packed(bufs), Packed(xs)
    A, X = Packed(bufs), Packed(xs)
    info = i32(np.full(len(ns), 99))
    plan = hip.DposvPlan(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, max(ns))
    wants = []
    for i, (n, l) in enumerate(zip(ns, ldas)):
        wa, wx = bufs[i].copy(), xs[i].copy()
        oracle.dposv(n, wa, l, wx)
        wants.append((wa, wx))
    for rep in range(4):
        A.dev.copy_(A0.dev)
        X.dev.copy_(X0.dev)
        info.fill_(99)
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        _ = (A.dev.sum() + X.dev.sum()).item()  # eager work between replays
        ga, gx = A.all(), X.all()
        assert np.all(info.cpu().numpy() == 0), rep
        for i in range(len(ns)):
            assert_bits_equal(ga[i], wants[i][0], f"replay {rep} A {i}")
            assert_bits_equal(gx[i], wants[i][1], f"replay {rep} x {i}")
    plan.close()


# ------------------------------------------------------------------ TRSV
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("diag", [131, 132])
def test_trsv_variants(uplo, trans, diag):
    hip = hipmod()
    rng = np.random.default_rng(uplo * 7 + trans + diag)
    ns = [1, 2, 3, 7, 16, 33, 64, 65, 100, 128] + [int(v) for v in rng.integers(1, 129, size=40)]
    ldas = [n + int(rng.integers(0, 4)) for n in ns]
    incs = [int(rng.integers(1, 4)) for _ in ns]
    bufs = []
    for i, (n, l) in enumerate(zip(ns, ldas)):
        a = rand_buf(500 + i, n, n, l)
        for j in range(n):
            a[j * l + j] = 2.0 + abs(a[j * l + j]) * np.sqrt(n)
        bufs.append(a)
    xs = [rand_buf(900 + i, 1, n * inc, 1) for i, (n, inc) in enumerate(zip(ns, incs))]
    A, X = Packed(bufs), Packed(xs)
    info = i32(np.full(len(ns), 99))
    hip.dtrsv_vbatched(uplo, trans, diag, i32(ns), A.ptrs, i32(ldas), X.ptrs, i32(incs), info, max(ns))
    torch.cuda.synchronize()
    gx = X.all()
    assert np.all(info.cpu().numpy() == 0)
    for i, (n, l, inc) in enumerate(zip(ns, ldas, incs)):
        want = xs[i].copy()
        if diag == 131:
            oracle.dtrsv(uplo, trans, n, bufs[i], l, want, inc)
        else:
            a = bufs[i].copy()
            for j in range(n):
                a[j * l + j] = 1.0  # LuInverse's unit-diagonal copy
            oracle.dtrsv(uplo, trans, n, a, l, want, inc)
        assert_bits_equal(gx[i], want, f"trsv {uplo}/{trans}/{diag} n={n} inc={inc}")


# ------------------------------------------------------------------ LU solve
def test_getrs_after_getrf():
    hip = hipmod()
    rng = np.random.default_rng(21)
    ns = [int(v) for v in rng.integers(1, 129, size=120)]
    cases = [(n, n, n + 1) for n in ns]
    bufs = [rand_buf(40 + i, n, n, n + 1, dist="normal") for i, n in enumerate(ns)]
    A = Packed(bufs)
    P = Packed([np.zeros(n, dtype=np.int32) for n in ns], dtype=torch.int32)
    sign = i32(np.zeros(len(ns)))
    info = i32(np.full(len(ns), 9))
    hip.dgetrf_vbatched(i32(ns), i32(ns), A.ptrs, i32([c[2] for c in cases]), P.ptrs, sign, info, 128, 128)
    bs = [rng.standard_normal(n) for n in ns]
    Bv = Packed(bs)
    Xv = Packed([np.zeros(n) for n in ns])
    info2 = i32(np.full(len(ns), 9))
    hip.dgetrs_vbatched(i32(ns), A.ptrs, i32([c[2] for c in cases]), P.ptrs, Bv.ptrs, Xv.ptrs, info2, 128)
    torch.cuda.synchronize()
    gx = Xv.all()
    assert np.all(info2.cpu().numpy() == 0)
    for i, n in enumerate(ns):
        lu = bufs[i].copy()
        piv, _ = oracle.dgetrf(n, n, lu, n + 1)
        want = oracle.dgetrs(n, lu, n + 1, piv, bs[i])
        assert_bits_equal(gx[i], want, f"getrs n={n}")


@pytest.mark.parametrize("nmax", [64, 128])
def test_getri_after_getrf_variable_sizes(nmax):
    """Factorization.Lu then Factorization.LuInverse on the device for 120
    matrices of order 1..nmax (lda = n + 1, NaN padding): the inverse
    bit-identical to the oracle; a matrix with an exactly zero pivot gets
    info = j + 1 and keeps its factors; a bad order is rejected alone.
    nmax = 64 runs the one-workgroup kernel, 128 the per-column solves."""
    hip = hipmod()
    rng = np.random.default_rng(31 + nmax)
    ns = [int(v) for v in rng.integers(1, nmax + 1, size=120)]
    ns[:3] = [1, 2, nmax]
    bufs = []
    for i, n in enumerate(ns):
        b = np.full((n + 1) * n, np.nan)
        col = np.random.default_rng(500 + i)
        for j in range(n):
            b[j * (n + 1):j * (n + 1) + n] = col.standard_normal(n)
        bufs.append(b)
    sing = 5  # a zero last column makes U[n-1, n-1] exactly zero
    n5 = ns[sing]
    bufs[sing][(n5 - 1) * (n5 + 1):(n5 - 1) * (n5 + 1) + n5] = 0.0
    ldas = [n + 1 for n in ns]
    A = Packed(bufs)
    P = Packed([np.zeros(n, dtype=np.int32) for n in ns], dtype=torch.int32)
    sign = i32(np.zeros(len(ns)))
    info = i32(np.full(len(ns), 9))
    hip.dgetrf_vbatched(i32(ns), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, nmax, nmax)
    torch.cuda.synchronize()
    lus = A.all()
    nbad = list(ns)
    nbad[7] = 0
    info2 = i32(np.full(len(ns), 9))
    hip.dgetri_vbatched(i32(nbad), A.ptrs, i32(ldas), P.ptrs, info2, nmax)
    torch.cuda.synchronize()
    got, gi = A.all(), info2.cpu().numpy()
    gp = P.all()
    for i, n in enumerate(ns):
        lu = bufs[i].copy()
        piv, _ = oracle.dgetrf(n, n, lu, n + 1)
        assert np.array_equal(gp[i], piv)
        assert_bits_equal(lus[i], lu, f"getrf n={n}")
        if i == 7:
            assert gi[i] == -3
            assert_bits_equal(got[i], lu, "rejected matrix unchanged")
            continue
        want = lu.copy()
        r = oracle.dgetri(n, want, n + 1, piv)
        assert gi[i] == r, (i, n, gi[i], r)
        assert_bits_equal(got[i], want, f"getri n={n}")
    assert gi[sing] == n5


def test_matflat_luinverse_on_device():
    from matflat_amd.matflat import ArgumentException, Factorization, MatrixFactorizationException
    n, lda = 10, 17
    a = rand_buf(42, n, n, lda)
    dev = torch.from_numpy(a.copy()).cuda()
    piv = torch.zeros(n, dtype=torch.int32, device="cuda")
    Factorization.Lu(n, n, dev, lda, piv)
    Factorization.LuInverse(n, dev, lda, piv)
    want = a.copy()
    wpiv, _ = oracle.dgetrf(n, n, want, lda)
    oracle.dgetri(n, want, lda, wpiv)
    assert_bits_equal(dev.cpu().numpy(), want, "LuInverse")
    z = torch.zeros(4, dtype=torch.float64, device="cuda")
    zp = torch.zeros(2, dtype=torch.int32, device="cuda")
    with pytest.raises(MatrixFactorizationException):
        Factorization.LuInverse(2, z, 2, zp)
    with pytest.raises(ArgumentException):
        Factorization.LuInverse(3, dev, 2, piv)


# ------------------------------------------------------------------ GEMM
def gemm_batch(complex_: bool, seed: int, count: int, shapes=None):
    rng = np.random.default_rng(seed)
    ops = [111, 112, 113, 114] if complex_ else [111, 112]
    recs = []
    for i in range(count):
        if shapes is not None and i < len(shapes):
            m, n, k = shapes[i][:3]
        else:
            m, n, k = (int(v) for v in rng.integers(1, 129, size=3))
        ta, tb = int(rng.choice(ops)), int(rng.choice(ops))
        ra, ca = (m, k) if ta in (111, 114) else (k, m)
        rb, cb = (k, n) if tb in (111, 114) else (n, k)
        lda, ldb, ldc = ra + int(rng.integers(0, 4)), rb + int(rng.integers(0, 4)), m + int(rng.integers(0, 4))
        gen = rand_cbuf if complex_ else rand_buf
        a, b, c = gen(3 * i, ra, ca, lda, "normal"), gen(3 * i + 1, rb, cb, ldb, "normal"), gen(3 * i + 2, m, n, ldc)
        if i % 3 == 0:
            alpha, beta = (1.0, 0.0)
        else:
            alpha, beta = rng.uniform(-1, 1), rng.uniform(-1, 1)
            if complex_:
                alpha, beta = complex(alpha, rng.uniform(-1, 1)), complex(beta, rng.uniform(-1, 1))
        recs.append((ta, tb, m, n, k, lda, ldb, ldc, a, b, c, alpha, beta))
    return recs


@pytest.mark.parametrize("complex_", [False, True])
def test_gemm_vbatched_bit_exact(complex_):
    hip = hipmod()
    recs = gemm_batch(complex_, 17, 400, shapes=[(m, n, k) for (m, n, k, *_r) in GEMM_CASES])
    dt = torch.complex128 if complex_ else torch.float64
    A = Packed([r[8] for r in recs], dtype=dt)
    Bm = Packed([r[9] for r in recs], dtype=dt)
    C = Packed([r[10] for r in recs], dtype=dt)
    if complex_:
        al = f64(np.array([[complex(r[11]).real, complex(r[11]).imag] for r in recs]).reshape(-1))
        be = f64(np.array([[complex(r[12]).real, complex(r[12]).imag] for r in recs]).reshape(-1))
    else:
        al = f64([r[11] for r in recs])
        be = f64([r[12] for r in recs])
    info = i32(np.full(len(recs), 99))
    col = lambda j: i32([r[j] for r in recs])  # noqa: E731
    hip.gemm_vbatched(col(0), col(1), col(2), col(3), col(4), al, A.ptrs, col(5), Bm.ptrs, col(6), be, C.ptrs,
                      col(7), info, 128, 128, 128, complex_=complex_)
    torch.cuda.synchronize()
    gc = C.all()
    assert np.all(info.cpu().numpy() == 0)
    for i, (ta, tb, m, n, k, lda, ldb, ldc, a, b, c, alpha, beta) in enumerate(recs):
        want = c.copy()
        fn = oracle.zgemm if complex_ else oracle.dgemm
        fn(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
        assert_bits_equal(gc[i], want, f"gemm {i} {(ta, tb, m, n, k)}")


def _op(buf, ld, t, rows, cols):
    """op(X) as a dense rows x cols array from a column-major buffer; t is the
    CBLAS transpose code (111 N, 112 T, 113 C, 114 conj without transpose)."""
    if t in (111, 114):
        x = view(buf, rows, cols, ld)
    else:
        x = view(buf, cols, rows, ld).T
    return np.conj(x) if t in (113, 114) else x


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("complex_", [False, True])
def test_gemm_plan_replays_match_eager(complex_, mode):
    """A batched GEMM captured into a plan and replayed four times, with
    eager work and synchronisation between replays (C restored before each),
    gives exactly the bits of the eager call in the same math mode."""
    hip = hipmod()
    recs = gemm_batch(complex_, 37, 200, shapes=[(m, n, k) for (m, n, k, *_r) in GEMM_CASES])
    dt = torch.complex128 if complex_ else torch.float64
    A = Packed([r[8] for r in recs], dtype=dt)
    Bm = Packed([r[9] for r in recs], dtype=dt)
    C0 = Packed([r[10] for r in recs], dtype=dt)
    C = Packed([r[10] for r in recs], dtype=dt)
    if complex_:
        al = f64(np.array([[complex(r[11]).real, complex(r[11]).imag] for r in recs]).reshape(-1))
        be = f64(np.array([[complex(r[12]).real, complex(r[12]).imag] for r in recs]).reshape(-1))
    else:
        al = f64([r[11] for r in recs])
        be = f64([r[12] for r in recs])
    col = lambda j: i32([r[j] for r in recs])  # noqa: E731
    args = (col(0), col(1), col(2), col(3), col(4), al, A.ptrs, col(5), Bm.ptrs, col(6), be, C.ptrs, col(7))
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST if mode == "fast" else hip.MATH_EXACT)
    info = i32(np.full(len(recs), 99))
    hip.gemm_vbatched(*args, info, 128, 128, 128, complex_=complex_, handle=h)
    torch.cuda.synchronize()
    want = C.all()
    assert np.all(info.cpu().numpy() == 0)
    plan = hip.GemmPlan(*args, info, 128, 128, 128, complex_=complex_, handle=h)
    for rep in range(4):
        C.dev.copy_(C0.dev)
        info.fill_(99)
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        _ = C.dev.abs().sum().item()  # eager work between replays
        assert np.all(info.cpu().numpy() == 0), rep
        got = C.all()
        for i in range(len(recs)):
            assert_bits_equal(got[i], want[i], f"replay {rep} product {i}")
    plan.close()
    h.close()


@pytest.mark.parametrize("complex_", [False, True])
def test_gemm_fast_mode_within_k_eps(complex_):
    """MFH_MATH_FAST runs the products on FP64 matrix cores; each product
    must satisfy ||C - C_ref||_F <= 4 k eps (|alpha| ||op(A)||_F ||op(B)||_F
    + |beta| ||C0||_F) against the oracle's C_ref (MatFlat's order).  Rows
    and columns outside m x n (the ld padding) must be untouched."""
    hip = hipmod()
    recs = gemm_batch(complex_, 29, 300, shapes=[(m, n, k) for (m, n, k, *_r) in GEMM_CASES])
    dt = torch.complex128 if complex_ else torch.float64
    A = Packed([r[8] for r in recs], dtype=dt)
    Bm = Packed([r[9] for r in recs], dtype=dt)
    C = Packed([r[10] for r in recs], dtype=dt)
    if complex_:
        al = f64(np.array([[complex(r[11]).real, complex(r[11]).imag] for r in recs]).reshape(-1))
        be = f64(np.array([[complex(r[12]).real, complex(r[12]).imag] for r in recs]).reshape(-1))
    else:
        al = f64([r[11] for r in recs])
        be = f64([r[12] for r in recs])
    info = i32(np.full(len(recs), 99))
    col = lambda j: i32([r[j] for r in recs])  # noqa: E731
    h = hip.Handle()
    h.set_math_mode(hip.MATH_FAST)
    hip.gemm_vbatched(col(0), col(1), col(2), col(3), col(4), al, A.ptrs, col(5), Bm.ptrs, col(6), be, C.ptrs,
                      col(7), info, 128, 128, 128, complex_=complex_, handle=h)
    torch.cuda.synchronize()
    h.close()
    gc = C.all()
    assert np.all(info.cpu().numpy() == 0)
    eps = np.finfo(np.float64).eps
    for i, (ta, tb, m, n, k, lda, ldb, ldc, a, b, c, alpha, beta) in enumerate(recs):
        want = c.copy()
        fn = oracle.zgemm if complex_ else oracle.dgemm
        fn(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
        got = gc[i]
        err = np.linalg.norm(view(got, m, n, ldc) - view(want, m, n, ldc))
        opa, opb = _op(a, lda, ta, m, k), _op(b, ldb, tb, k, n)
        scale = abs(alpha) * np.linalg.norm(opa) * np.linalg.norm(opb) + abs(beta) * np.linalg.norm(view(c, m, n, ldc))
        assert err <= 4 * k * eps * scale, (i, ta, tb, m, n, k, err, scale)
        # padding rows of C are not written
        for j in range(n):
            assert_bits_equal(got[j * ldc + m:(j + 1) * ldc], c[j * ldc + m:(j + 1) * ldc], f"pad {i}")


def test_gemm_aliasing_rejected_per_product():
    hip = hipmod()
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = torch.tensor([buf.data_ptr()], dtype=torch.int64, device="cuda")
    info = i32([99])
    one = f64([1.0])
    hip.gemm_vbatched(i32([111]), i32([111]), i32([4]), i32([4]), i32([4]), one, p, i32([4]), p, i32([4]),
                      f64([0.0]), p, i32([4]), info, 4, 4, 4)
    torch.cuda.synchronize()
    assert int(info.item()) == -1


# ------------------------------------------------------------------ adjoint
def test_posv_backward_matches_oracle():
    hip = hipmod()
    rng = np.random.default_rng(31)
    ns = [1, 2, 5, 32, 64, 100, 128] + [int(v) for v in rng.integers(1, 129, size=60)]
    ls, xs, ws = [], [], []
    for i, n in enumerate(ns):
        buf = spd_upper_buf(600 + i, n, n)
        x = rng.standard_normal(n)
        assert oracle.dposv(n, buf, n, x) == 0
        ls.append(buf)
        xs.append(x)
        ws.append(rng.standard_normal(n))
    L, X, W = Packed(ls), Packed(xs), Packed(ws)
    AB = Packed([np.zeros(n * n) for n in ns])
    BB = Packed([np.zeros(n) for n in ns])
    info = i32(np.full(len(ns), 99))
    hip.dposv_backward_vbatched(i32(ns), L.ptrs, i32(ns), X.ptrs, W.ptrs, AB.ptrs, i32(ns), BB.ptrs, info, 128)
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    gab, gbb = AB.all(), BB.all()
    for i, n in enumerate(ns):
        abar, bbar = oracle.dposv_backward(n, ls[i], n, xs[i], ws[i])
        assert_bits_equal(gbb[i], bbar, f"bbar n={n}")
        assert_bits_equal(gab[i], abar, f"abar n={n}")


# ------------------------------------------------------------------ MatFlat signatures
def test_matflat_signatures_on_device():
    from matflat_amd.matflat import (ArgumentException, Blas, Factorization, MatrixFactorizationException,
                                     Transpose, Uplo)
    m, n, lda = 23, 11, 31
    a = rand_buf(42, m, n, lda)
    dev = torch.from_numpy(a.copy()).cuda()
    piv = torch.zeros(m, dtype=torch.int32, device="cuda")
    sign = Factorization.Lu(m, n, dev, lda, piv)
    want = a.copy()
    wpiv, wsign = oracle.dgetrf(m, n, want, lda)
    assert_bits_equal(dev.cpu().numpy(), want, "Lu")
    assert np.array_equal(piv.cpu().numpy(), wpiv) and sign == wsign
    with pytest.raises(ArgumentException):
        Factorization.Lu(0, 3, dev, 3, piv)
    with pytest.raises(ArgumentException):
        Factorization.Lu(4, 3, dev, 3, piv)
    bad = torch.tensor([1.0, np.nan, 2.0, -5.0], dtype=torch.float64, device="cuda")
    with pytest.raises(MatrixFactorizationException):
        Factorization.Cholesky(2, bad, 2)
    x = torch.ones(5, dtype=torch.float64, device="cuda")
    tri = torch.from_numpy(np.eye(5).reshape(-1) * 2.0).cuda()
    Blas.SolveTriangular(Uplo.Lower, Transpose.NoTrans, 5, tri, 5, x, 1)
    assert torch.all(x == 0.5)
    c = torch.zeros(4, dtype=torch.float64, device="cuda")
    aa = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64, device="cuda")
    Blas.MulMatMat(Transpose.NoTrans, Transpose.NoTrans, 2, 2, 2, aa, 2, aa, 2, c, 2)
    assert c.cpu().tolist() == [7.0, 10.0, 15.0, 22.0]
    with pytest.raises(ArgumentException):
        Blas.MulMatMat(Transpose.NoTrans, Transpose.NoTrans, 2, 2, 2, aa, 2, aa, 2, aa, 2)


def test_matflat_dropin_ignores_fast_default_handle():
    """The drop-in calls run in exact mode whatever the default handle's
    mode: a Cholesky that fails keeps MatFlat's partial writes (FAST mode
    would leave the matrix unchanged) and an LU stays bit-identical."""
    from matflat_amd import hip
    from matflat_amd.matflat import Factorization, MatrixFactorizationException
    dh = hip.default_handle()
    old = dh.math_mode()
    dh.set_math_mode(hip.MATH_FAST)
    try:
        n = 6
        rng = np.random.default_rng(7)
        b = rng.standard_normal((n, n))
        s = b @ b.T + n * np.eye(n)
        s[4, 4] = -50.0  # row 4 fails after rows 0..3 are written
        a = np.asfortranarray(s).reshape(-1, order="F").copy()
        want = a.copy()
        assert oracle.dpotrf(n, want, n) == 5
        dev = torch.from_numpy(a.copy()).cuda()
        with pytest.raises(MatrixFactorizationException):
            Factorization.Cholesky(n, dev, n)
        assert_bits_equal(dev.cpu().numpy(), want, "partial writes")
        m = 40
        la = rand_buf(11, m, m, m)
        dev = torch.from_numpy(la.copy()).cuda()
        piv = torch.zeros(m, dtype=torch.int32, device="cuda")
        sign = Factorization.Lu(m, m, dev, m, piv)
        wpiv, wsign = oracle.dgetrf(m, m, la, m)
        assert_bits_equal(dev.cpu().numpy(), la, "Lu")
        assert np.array_equal(piv.cpu().numpy(), wpiv) and sign == wsign
    finally:
        dh.set_math_mode(old)


# ------------------------------------------------------------------ GEMV
@pytest.mark.parametrize("seed", [0, 1])
def test_gemv_variable_sizes_bit_exact(seed):
    """A batch of 300 products with m, n ~ U{1..300}, random transposes,
    leading dimensions and strides: y bit-identical to the oracle."""
    hip = hipmod()
    rng = np.random.default_rng(4000 + seed)
    cnt = 300
    tr = rng.choice([111, 112], cnt)
    m = rng.integers(1, 301, cnt)
    n = rng.integers(1, 301, cnt)
    lda = m + rng.integers(0, 4, cnt)
    incx = rng.integers(1, 4, cnt)
    incy = rng.integers(1, 4, cnt)
    xl = np.where(tr == 112, m, n)
    yl = np.where(tr == 112, n, m)
    As = [rng.standard_normal(int(l * c)) for l, c in zip(lda, n)]
    Xs = [rng.standard_normal(int(l * c)) for l, c in zip(xl, incx)]
    Ys = [rng.standard_normal(int(l * c)) for l, c in zip(yl, incy)]
    A, X, Y = Packed(As), Packed(Xs), Packed([y.copy() for y in Ys])
    info = i32(np.full(cnt, 99))
    hip.dgemv_vbatched(i32(tr), i32(m), i32(n), A.ptrs, i32(lda), X.ptrs, i32(incx), Y.ptrs, i32(incy), info,
                       int(m.max()), int(n.max()))
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    gy = Y.all()
    for i in range(cnt):
        want = Ys[i].copy()
        oracle.dgemv(int(tr[i]), int(m[i]), int(n[i]), As[i], int(lda[i]), Xs[i], int(incx[i]), want, int(incy[i]))
        assert_bits_equal(gy[i], want, f"gemv {i}")


def test_gemv_argument_checks():
    """Per-product checks in Blas.MulMatVec.cs order, the transpose last:
    info = -(position in mfh_dgemv_vbatched's argument list), -1 for y
    aliasing a or x; the other products of the batch still run."""
    hip = hipmod()
    from matflat_amd.hip import ArgumentError
    good = np.arange(1.0, 7.0)  # 2 x 3, lda 2
    cases = [  # (trans, m, n, lda, incx, incy, y aliases a) -> info
        ((111, 2, 3, 2, 1, 1, False), 0),
        ((111, 0, 3, 2, 1, 1, False), -4),
        ((111, 2, 0, 2, 1, 1, False), -5),
        ((111, 2, 3, 1, 1, 1, False), -7),
        ((111, 2, 3, 2, 0, 1, False), -9),
        ((111, 2, 3, 2, 1, 0, False), -11),
        ((111, 2, 3, 2, 1, 1, True), -1),
        ((113, 2, 3, 2, 1, 1, False), -3),
        ((113, 0, 3, 2, 1, 1, False), -4),  # the transpose is checked last
        ((112, 2, 3, 2, 1, 1, False), 0),
    ]
    A = Packed([good.copy() for _ in cases])
    X = Packed([np.ones(3) for _ in cases])
    Y = Packed([np.full(3, 7.0) for _ in cases])
    yp = Y.ptrs.clone()
    for i, (c, _) in enumerate(cases):
        if c[6]:
            yp[i] = A.ptrs[i]
    col = lambda j: i32([c[j] for c, _ in cases])  # noqa: E731
    info = i32(np.full(len(cases), 99))
    hip.dgemv_vbatched(col(0), col(1), col(2), A.ptrs, col(3), X.ptrs, col(4), yp, col(5), info, 2, 3)
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [w for _, w in cases]
    gy = Y.all()
    assert gy[0][:2].tolist() == [9.0, 12.0]  # rows of [[1, 3, 5], [2, 4, 6]] summed
    assert gy[9].tolist() == [3.0, 7.0, 11.0]  # columns summed
    for i in range(1, 9):
        assert np.all(gy[i] == 7.0)
    with pytest.raises(ArgumentError):
        hip.dgemv_vbatched(col(0), col(1), col(2), None, col(3), X.ptrs, col(4), Y.ptrs, col(5), info, 2, 3)


def test_matflat_mulmatvec_on_device():
    from matflat_amd.matflat import ArgumentException, Blas, Transpose
    m, n, lda = 23, 11, 31
    a = rand_buf(42, m, n, lda)
    x = np.random.default_rng(5).standard_normal(2 * n)
    y = np.full(3 * m, 0.25)
    da, dx, dy = (torch.from_numpy(v.copy()).cuda() for v in (a, x, y))
    Blas.MulMatVec(Transpose.NoTrans, m, n, da, lda, dx, 2, dy, 3)
    want = y.copy()
    oracle.dgemv(111, m, n, a, lda, x, 2, want, 3)
    assert_bits_equal(dy.cpu().numpy(), want, "MulMatVec NoTrans")
    with pytest.raises(ArgumentException):
        Blas.MulMatVec(Transpose.NoTrans, m, n, da, lda, dx, 2, da, 3)
    with pytest.raises(ArgumentException):
        Blas.MulMatVec(Transpose.ConjTrans, m, n, da, lda, dx, 2, dy, 3)
    with pytest.raises(ArgumentException):
        Blas.MulMatVec(Transpose.NoTrans, m, n, da, m - 1, dx, 2, dy, 3)


# ------------------------------------------------------------------ FAST math mode
@pytest.mark.parametrize("waves", [0, 1, 2, 4])
@pytest.mark.parametrize("posv", [False, True])
def test_fast_mode_cholesky_residuals(posv, waves):
    """FAST mode (FMA, right-looking) agrees with MatFlat to rounding: the
    dpot01 residual ratio ||L L^T - A||_1 / (n ||A||_1 eps) stays <= 10,
    L matches the oracle within 1e-12 relative, the strict upper triangle is
    exactly zero, padding is untouched, and the solve has backward error
    <= n*eps."""
    from helpers import backward_error, chol_residual_ratio, view
    hip = hipmod()
    ns, ldas = chol_cases(21, 300, nmax=128)
    ns[:6] = [1, 2, 8, 9, 100, 127]
    ldas[:6] = [1, 3, 8, 12, 104, 127]
    bufs = [spd_upper_buf(700 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    xs = [np.random.default_rng(900 + i).standard_normal(n) for i, n in enumerate(ns)]
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST)
    assert h.math_mode() == hip.MATH_FAST
    if waves:
        for rb in range(2 if waves <= 2 else 4, 18):
            h.set_fast_waves(rb, waves)
            assert h.fast_waves(rb) == waves
    A, X = Packed(bufs), Packed(xs)
    info = i32(np.full(len(ns), 99))
    if posv:
        hip.dposv_vbatched(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, max(ns), handle=h)
    else:
        hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=h)
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    ga, gx = A.all(), X.all()
    eps = np.finfo(np.float64).eps
    for i, (n, l) in enumerate(zip(ns, ldas)):
        full = view(bufs[i], n, n, l)
        a_sym = np.triu(full) + np.triu(full, 1).T
        lg = view(ga[i], n, n, l)
        assert np.all(np.triu(lg, 1) == 0.0)
        assert chol_residual_ratio(a_sym, lg) <= 10.0, (i, n)
        want = bufs[i].copy()
        wx = xs[i].copy()
        oracle.dposv(n, want, l, wx)
        lw = view(want, n, n, l)
        assert np.max(np.abs(lg - lw)) <= 1e-12 * np.max(np.abs(lw)) * max(1, n), (i, n)
        for j in range(n):
            assert np.all(np.isnan(ga[i][j * l + n:(j + 1) * l]))
        if posv:
            assert backward_error(a_sym, gx[i], xs[i]) <= n * eps, (i, n)
    h.close()


@pytest.mark.parametrize("use_plan", [False, True])
def test_fast_mode_bucket_streams_same_bits(use_plan):
    """The FAST size buckets forked over 1..4 streams (eagerly, and inside a
    captured plan replayed several times with eager work in between) give
    the same bits: the fork only changes when each bucket runs.  The
    one-stream result is also checked against the oracle by residual."""
    from helpers import backward_error, view
    hip = hipmod()
    ns, ldas = chol_cases(41, 600, nmax=100)
    ns[:4] = [1, 9, 99, 100]
    ldas[:4] = [1, 16, 104, 100]
    bufs = [spd_upper_buf(2100 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    xs = [np.random.default_rng(2900 + i).standard_normal(n) for i, n in enumerate(ns)]
    A0, X0 = Packed(bufs), Packed(xs)
    ref = None
    for streams in (1, 2, 3, 4):
        h = hip.Handle()
        h.bind_stream()
        h.set_math_mode(hip.MATH_FAST)
        h.set_bucket_streams(streams)
        assert h.bucket_streams() == streams
        A, X = Packed(bufs), Packed(xs)
        info = i32(np.full(len(ns), 99))
        if use_plan:
            plan = hip.DposvPlan(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, max(ns), handle=h)
            reps = 3
        else:
            reps = 1
        for rep in range(reps):
            A.dev.copy_(A0.dev)
            X.dev.copy_(X0.dev)
            info.fill_(99)
            torch.cuda.synchronize()
            if use_plan:
                plan.execute()
            else:
                hip.dposv_vbatched(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, max(ns), handle=h)
            torch.cuda.synchronize()
            _ = (A.dev.sum() + X.dev.sum()).item()  # eager work between replays
            assert np.all(info.cpu().numpy() == 0), (streams, rep)
            got = (A.all(), X.all())
            if ref is None:
                ref = got
                eps = np.finfo(np.float64).eps
                for i, (n, l) in enumerate(zip(ns, ldas)):
                    full = view(bufs[i], n, n, l)
                    a_sym = np.triu(full) + np.triu(full, 1).T
                    assert backward_error(a_sym, got[1][i], xs[i]) <= n * eps, (i, n)
            for i in range(len(ns)):
                assert_bits_equal(got[0][i], ref[0][i], f"streams {streams} rep {rep} A {i}")
                assert_bits_equal(got[1][i], ref[1][i], f"streams {streams} rep {rep} x {i}")
        if use_plan:
            plan.close()
        h.close()
    h = hip.Handle()
    from matflat_amd.hip import ArgumentError
    with pytest.raises(ArgumentError):
        h.set_bucket_streams(0)
    with pytest.raises(ArgumentError):
        h.set_bucket_streams(5)
    h.close()


@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("posv", [False, True])
def test_fast_mode_not_positive_definite(waves, posv):
    """FAST mode reports the failing pivot like MatFlat (info = j+1 for a
    clearly indefinite leading block) and stops there: the columns of the
    8-column panels before the failing one hold L (equal to the oracle's
    partial factor on the rows MatFlat had finished, within rounding), the
    right-hand side is left unchanged, and the neighbours in the batch are
    factored normally."""
    from helpers import view
    hip = hipmod()
    ns = [5, 40, 70, 90, 12, 33]
    bad = {1: 17, 3: 60, 5: 7}  # matrix index -> row whose diagonal is made negative
    bufs = []
    for i, n in enumerate(ns):
        b = spd_upper_buf(1300 + i, n, n)
        if i in bad:
            j = bad[i]
            b[j + j * n] = -1.0e3 * n
        bufs.append(b)
    xs = [np.random.default_rng(1400 + i).standard_normal(n) for i, n in enumerate(ns)]
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST)
    for rb in range(2, 18):
        if waves <= 2 or rb >= 4:
            h.set_fast_waves(rb, waves)
    A, X = Packed(bufs), Packed(xs)
    info = i32(np.full(len(ns), 99))
    if posv:
        hip.dposv_vbatched(i32(ns), A.ptrs, i32(ns), X.ptrs, info, max(ns), handle=h)
    else:
        hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ns), info, max(ns), handle=h)
    torch.cuda.synchronize()
    got = info.cpu().numpy()
    ga, gx = A.all(), X.all()
    for i, n in enumerate(ns):
        if i in bad:
            j = bad[i]
            assert got[i] == j + 1, (i, got)
            want = bufs[i].copy()
            assert oracle.dpotrf(n, want, n) == j + 1
            c = 8 * (j // 8)  # columns of the panels before the failing one
            lg, lw = view(ga[i], n, n, n), view(want, n, n, n)
            if c:
                sub_g, sub_w = np.tril(lg[:j, :c]), np.tril(lw[:j, :c])
                assert np.max(np.abs(sub_g - sub_w)) <= 1e-12 * np.max(np.abs(sub_w)) * n, i
            assert_bits_equal(gx[i], xs[i], f"x of failed matrix {i}")
        else:
            assert got[i] == 0
    h.close()
