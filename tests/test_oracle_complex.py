"""Pins the oracle's complex128 LU, Cholesky and triangular solve
(oracle/matflat_oracle.c: mfo_zdiv, mfo_zdot, mfo_zgetrf, mfo_zpotrf,
mfo_ztrsv) before the GPU kernels are checked against it.

* Division: BCL Complex division is Smith's method with the branch on
  |d| < |c|.  CPython's complex division is Smith's method too, with the
  branch on |c| >= |d|; at |c| == |d| both branches perform the same exact
  operations, so the two agree bit for bit on all finite operands (checked,
  with hand-derived cases).
* Dot order: hand-derived complex cases whose results differ for every other
  pairing.
* LU: LAPACK zgetrf (OpenBLAS) pivots on |re| + |im| (izamax) with the first
  maximum, exactly MatFlat's FastMagnitude rule (Internals.cs:15-19), so the
  permutation matches and the factors agree to 1e-12 (LuTests.cs:143's
  tolerance); exact rational LU on small Gaussian-integer matrices.
* Cholesky: LAPACK zpotrf (lower, A = L L^H) at 1e-12, the 1x1 and 2x2 closed
  forms, the failure row and the partial writes (rows before it equal the
  factor of the leading block, everything else untouched).
* Triangular solve: LAPACK ztrtrs for NoTrans, Trans and ConjTrans;
  ConjNoTrans through conj(A) x = b <=> A conj(x) = conj(b).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import oracle
from helpers import EPS, hpd_upper_cbuf, ipiv_to_perm, rand_cbuf, view
from test_oracle import LU_CASES

NOTRANS, TRANS, CONJTRANS, CONJNOTRANS = 111, 112, 113, 114
UPPER, LOWER = 121, 122


def bits(z: complex):
    return np.array([z.real, z.imag]).view(np.uint64).tolist()


# ------------------------------------------------------------ division
def test_zdiv_matches_smith_when_abs_differ():
    rng = np.random.default_rng(3)
    vals = rng.standard_normal((4000, 4)) * np.exp(rng.uniform(-20, 20, (4000, 4)))
    for a, b, c, d in vals:
        if abs(c) == abs(d):
            continue
        assert bits(oracle.zdiv(complex(a, b), complex(c, d))) == bits(complex(a, b) / complex(c, d))


def test_zdiv_equal_magnitudes():
    # |c| == |d|: r = c/d = +-1 exactly, so both of Smith's branches reduce to
    # the same exact operations ((a +- b) / 2c up to exact sign flips) and the
    # branch taken at equality cannot change a bit.
    # (1 + 2i) / (1 + i) = (3 + i) / 2
    assert oracle.zdiv(1 + 2j, 1 + 1j) == 1.5 + 0.5j
    # (1 + 2i) / (1 - i) = (1 + 2i)(1 + i) / 2 = (-1 + 3i) / 2
    assert oracle.zdiv(1 + 2j, 1 - 1j) == -0.5 + 1.5j
    rng = np.random.default_rng(11)
    for _ in range(2000):
        a, b, c = rng.standard_normal(3)
        d = c if rng.random() < 0.5 else -c
        assert bits(oracle.zdiv(complex(a, b), complex(c, d))) == bits(complex(a, b) / complex(c, d))


def test_zdiv_by_real_matches_component_division():
    # Cholesky divides by (L[k,k], 0): |0| < |c| gives r = 0, den = c,
    # ((a + b*0)/c, (b - a*0)/c) = (a/c, b/c) for finite a, b
    rng = np.random.default_rng(5)
    for a, b, c in rng.standard_normal((500, 3)):
        c = abs(c) + 0.1
        assert oracle.zdiv(complex(a, b), complex(c, 0.0)) == complex(a / c, b / c)


# ------------------------------------------------------------ dot order
def test_zdot_order_odd_and_even():
    big = 2.0**60
    # odd n = 3, real parts of products [1, big, -big]: s = p0; s + (p1 + p2) = 1
    x = np.array([1, big, -big], dtype=np.complex128)
    assert oracle.zdot(0, 3, x, 1, np.ones(3, dtype=np.complex128), 1) == 1.0
    # even n = 4, products [big, 1, -big, 1]: (big + 1) + (-big + 1) = 0
    # (left to right: 1; odd-style: big + ... )
    x = np.array([big, 1, -big, 1], dtype=np.complex128)
    assert oracle.zdot(0, 4, x, 1, np.ones(4, dtype=np.complex128), 1) == 0.0
    # imaginary parts follow the same pairing: x = i*[big, 1, -big, 1]
    assert oracle.zdot(0, 4, 1j * x, 1, np.ones(4, dtype=np.complex128), 1) == 0.0


def test_zdot_conj_and_strides():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    y = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    for conj in (0, 1):
        for n, incx, incy in [(0, 1, 1), (1, 1, 1), (7, 3, 2), (10, 2, 4), (13, 1, 3)]:
            xv, yv = x[: max(n, 1) * incx:incx][:n], y[: max(n, 1) * incy:incy][:n]
            want = np.sum((np.conj(xv) if conj else xv) * yv)
            got = oracle.zdot(conj, n, x, incx, y, incy)
            assert abs(got - want) <= 1e-13 * (1 + np.sum(np.abs(xv) * np.abs(yv)))
    # conj(x)*y for a single product, exactly: (1 + 2i)^* (3 + 4i) = 11 - 2i
    assert oracle.zdot(1, 1, np.array([1 + 2j]), 1, np.array([3 + 4j]), 1) == 11 - 2j


# ------------------------------------------------------------ LU
@pytest.mark.parametrize("m,n,lda", LU_CASES)
def test_zlu_matches_openblas(m, n, lda):
    a = rand_cbuf(42, m, n, lda)
    a0 = a.copy()
    piv, sign = oracle.zgetrf(m, n, a, lda)
    lu_ref, ipiv, info = lapack.zgetrf(view(a0, m, n, lda))
    np.testing.assert_allclose(view(a, m, n, lda), lu_ref, rtol=0, atol=1e-12)
    assert np.array_equal(piv, ipiv_to_perm(ipiv, m))
    assert sign == (-1) ** int(np.sum(ipiv != np.arange(len(ipiv))))
    for j in range(n):
        assert np.array_equal(a[j * lda + m:(j + 1) * lda], a0[j * lda + m:(j + 1) * lda])


@pytest.mark.parametrize("n", [16, 64, 100])
def test_zlu_residual_and_bounds(n):
    a = rand_cbuf(3 + n, n, n, n, dist="normal")
    a0 = view(a, n, n, n)
    piv, _ = oracle.zgetrf(n, n, a, n)
    lu = view(a, n, n, n)
    assert sorted(piv.tolist()) == list(range(n))
    low = np.tril(lu, -1)
    # pivoting on |re| + |im| bounds the multipliers' modulus by sqrt(2)
    # (|a| <= |a|_1 <= |p|_1 <= sqrt(2) |p|)
    assert np.all(np.abs(low) <= np.sqrt(2.0) * (1 + 1e-15))
    l = low + np.eye(n)
    u = np.triu(lu)
    r = np.linalg.norm(a0[piv, :] - l @ u, 1) / (n * np.linalg.norm(a0, 1) * EPS)
    assert r <= 10.0


def _gauss_lu(mat):
    """Partial-pivot LU over Gaussian rationals (pairs of Fractions); pivot
    = first maximum of |re| + |im|."""
    m, n = len(mat), len(mat[0])
    a = [[(Fraction(v.real), Fraction(v.imag)) for v in row] for row in mat]
    perm = list(range(m))

    def mag(z):
        return abs(z[0]) + abs(z[1])

    def div(x, y):
        den = y[0] * y[0] + y[1] * y[1]
        return ((x[0] * y[0] + x[1] * y[1]) / den, (x[1] * y[0] - x[0] * y[1]) / den)

    def mul(x, y):
        return (x[0] * y[0] - x[1] * y[1], x[0] * y[1] + x[1] * y[0])

    for j in range(min(m, n)):
        p = max(range(j, m), key=lambda i: (mag(a[i][j]), -i))
        if p != j:
            a[p], a[j] = a[j], a[p]
            perm[p], perm[j] = perm[j], perm[p]
        if a[j][j] != (0, 0):
            for i in range(j + 1, m):
                f = div(a[i][j], a[j][j])
                a[i][j] = f
                for c in range(j + 1, n):
                    t = mul(f, a[j][c])
                    a[i][c] = (a[i][c][0] - t[0], a[i][c][1] - t[1])
    return perm, a


@pytest.mark.parametrize("seed", range(12))
def test_zlu_pivots_match_exact_rational(seed):
    rng = np.random.default_rng(100 + seed)
    m, n = int(rng.integers(2, 6)), int(rng.integers(2, 6))
    mat = rng.integers(-4, 5, (m, n)) + 1j * rng.integers(-4, 5, (m, n))
    a = np.asfortranarray(mat.astype(np.complex128)).reshape(-1, order="F").copy()
    piv, _ = oracle.zgetrf(m, n, a, m)
    perm, exact = _gauss_lu(mat.tolist())
    assert piv.tolist() == perm
    got = view(a, m, n, m)
    for i in range(m):
        for j in range(n):
            re, im = exact[i][j]
            assert abs(got[i, j] - complex(float(re), float(im))) <= 1e-12 * (1 + abs(got[i, j]))


def test_zlu_zero_matrix_and_ties():
    a = np.zeros(9, dtype=np.complex128)
    piv, sign = oracle.zgetrf(3, 3, a, 3)
    assert piv.tolist() == [0, 1, 2] and sign == 1 and np.all(a == 0)
    # column 0 = [1 + 2i, 3, -2 - 1i]: magnitudes 3, 3, 3 -> first maximum (row 0)
    a = np.array([1 + 2j, 3, -2 - 1j, 0, 1, 0, 0, 0, 1], dtype=np.complex128)
    piv, sign = oracle.zgetrf(3, 3, a, 3)
    assert piv[0] == 0 and sign == 1
    # a NaN below the diagonal is never chosen; NaN on the diagonal stays
    a = np.array([1, complex(np.nan, 0), 2j, 0, 1, 0, 0, 0, 1], dtype=np.complex128)
    piv, _ = oracle.zgetrf(3, 3, a, 3)
    assert piv[0] == 2
    a = np.array([complex(0, np.nan), 5, 7, 0, 1, 0, 0, 0, 1], dtype=np.complex128)
    piv, _ = oracle.zgetrf(3, 3, a, 3)
    assert piv[0] == 0


# ------------------------------------------------------------ Cholesky
@pytest.mark.parametrize("n,lda", [(1, 1), (2, 3), (3, 3), (11, 16), (23, 24), (33, 40), (64, 64)])
def test_zcholesky_matches_openblas(n, lda):
    a = hpd_upper_cbuf(n, n, lda)
    a0 = a.copy()
    assert oracle.zpotrf(n, a, lda) == 0
    full = view(a0, n, n, lda)
    herm = np.triu(full) + np.triu(full, 1).conj().T
    l_ref, info = lapack.zpotrf(herm, lower=1)
    assert info == 0
    got = view(a, n, n, lda)
    np.testing.assert_allclose(got, np.tril(l_ref), rtol=0, atol=1e-12 * max(1.0, np.abs(l_ref).max()))
    assert np.all(np.triu(got, 1) == 0)
    assert np.all(np.diag(got).imag == 0)
    for j in range(n):  # padding rows untouched
        assert np.array_equal(a[j * lda + n:(j + 1) * lda], a0[j * lda + n:(j + 1) * lda], equal_nan=True)


def test_zcholesky_closed_forms():
    a = np.array([4 + 0j])
    assert oracle.zpotrf(1, a, 1) == 0 and a[0] == 2
    # [[4, 2 + 2i], [2 - 2i, 6]]: L = [[2, 0], [1 - i, 2]]
    a = np.array([4, np.nan, 2 + 2j, 6], dtype=np.complex128)
    assert oracle.zpotrf(2, a, 2) == 0
    assert a.tolist() == [2, 1 - 1j, 0, 2]


def test_zcholesky_failure_partial_writes():
    n, f = 7, 4
    a = hpd_upper_cbuf(5, n, n)
    a[f * n + f] = -100.0  # row f fails
    a0 = a.copy()
    assert oracle.zpotrf(n, a, n) == f + 1
    full0 = view(a0, n, n, n)
    herm = np.triu(full0) + np.triu(full0, 1).conj().T
    l_lead, info = lapack.zpotrf(herm[:f, :f], lower=1)
    assert info == 0
    got = view(a, n, n, n)
    np.testing.assert_allclose(np.tril(got[:f, :f]), np.tril(l_lead), atol=1e-12 * np.abs(l_lead).max())
    # row f: L[f, 0:f] = conj(L_lead^-1 A[0:f, f])
    want_row = np.conj(np.linalg.solve(np.tril(l_lead), full0[:f, f]))
    np.testing.assert_allclose(got[f, :f], want_row, atol=1e-12 * np.abs(want_row).max())
    # untouched: upper triangle (diagonal of rows >= f included), rows > f of the lower triangle
    for j in range(n):
        for i in range(n):
            written = (i < f and j <= i) or (i == f and j < f)
            if not written:
                assert np.array_equal(a[i + j * n], a0[i + j * n], equal_nan=True), (i, j)


def test_zcholesky_not_hermitian_diag_imag_ignored():
    # only re(A[j,j]) is read (Factorization.Cholesky.cs:229)
    a = np.array([4 + 7j], dtype=np.complex128)
    assert oracle.zpotrf(1, a, 1) == 0 and a[0] == 2


# ------------------------------------------------------------ triangular solve
def _tri(seed, n, lda, lower):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    m = np.tril(m) if lower else np.triu(m)
    m[np.diag_indices(n)] += 4.0 * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    buf = rand_cbuf(seed + 1, lda, n, lda)  # garbage outside the triangle
    for j in range(n):
        for i in range(n):
            if (i >= j) if lower else (i <= j):
                buf[i + j * lda] = m[i, j]
    return buf, m


@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NOTRANS, TRANS, CONJTRANS, CONJNOTRANS])
@pytest.mark.parametrize("n,lda,incx", [(1, 1, 1), (5, 7, 2), (16, 16, 1), (33, 40, 3)])
def test_ztrsv_matches_openblas(uplo, trans, n, lda, incx):
    buf, m = _tri(n + lda + incx, n, lda, uplo == LOWER)
    rng = np.random.default_rng(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = np.full(n * incx, 99 + 99j, dtype=np.complex128)
    x[::incx] = b
    assert oracle.ztrsv(uplo, trans, n, buf, lda, x, incx) == 0
    lower = 1 if uplo == LOWER else 0
    if trans == CONJNOTRANS:
        want = np.conj(lapack.ztrtrs(m, np.conj(b), lower=lower, trans=0)[0])
    else:
        want = lapack.ztrtrs(m, b, lower=lower, trans={NOTRANS: 0, TRANS: 1, CONJTRANS: 2}[trans])[0]
    np.testing.assert_allclose(x[::incx], want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))
    if incx > 1:
        assert np.all(np.delete(x, np.s_[::incx]) == 99 + 99j)


def test_ztrsv_rejects_bad_enums():
    a = np.ones(4, dtype=np.complex128)
    x = np.ones(2, dtype=np.complex128)
    assert oracle.ztrsv(UPPER, 110, 2, a, 2, x, 1) == -2
    assert oracle.ztrsv(120, NOTRANS, 2, a, 2, x, 1) == -1
