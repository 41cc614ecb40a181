"""Pins the CPU oracle (oracle/matflat_oracle.c) before it is trusted.

No golden vectors exist in the reference (its tests take expected values
from OpenBLAS at run time).  The oracle is therefore pinned by:

* agreement with OpenBLAS through scipy at the reference tests' tolerances
  (LuTests.cs:143, CholeskyTests.cs:124, BlasTests.SolveTriangularDouble.cs:54,
  BlasTests.MulMatMatDouble.cs:51, BlasTests.MulMatMatComplex.cs:59-60);
* pivot vectors equal to LAPACK's swap sequence converted to a permutation;
* residual ratios (LAPACK dget01/dpot01) and structural invariants;
* hand-derived summation-order cases whose results differ for every other
  plausible pairing (they fail if a term, the odd/even split or the pairing
  were wrong);
* brute force: exact rational LU with partial pivoting on tiny integer
  matrices (fractions.Fraction), and the pivot rule on NaN/inf/tie columns;
* finite differences for the adjoint.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.linalg.lapack as lapack

import oracle
from helpers import (EPS, chol_residual_ratio, ipiv_to_perm, lu_residual_ratio, rand_buf, rand_cbuf,
                     spd_upper_buf, store, view)

LU_CASES = [(1, 1, 1), (1, 1, 3), (2, 2, 2), (2, 2, 3), (3, 3, 3), (3, 3, 5), (3, 1, 3), (3, 1, 5),
            (1, 3, 1), (1, 3, 3), (3, 4, 3), (3, 4, 5), (4, 3, 4), (4, 3, 5), (11, 23, 11),
            (11, 23, 17), (16, 8, 16), (16, 8, 32), (23, 11, 23), (23, 11, 31)]  # LuTests.cs:103-122


# ---------------------------------------------------------------- Dot order
def test_dot_order_odd_length():
    # n = 5, products p = [1, 2^60, -2^60, 1, 1].  MatFlat: s = p0;
    # s += (p1 + p2); s += (p3 + p4)  ->  1 + 0 + 2 = 3.
    # Left-to-right gives 2, even-style pairing gives 1.
    x = np.array([1.0, 2.0**60, -(2.0**60), 1.0, 1.0])
    y = np.ones(5)
    assert oracle.ddot(5, x, 1, y, 1) == 3.0


def test_dot_order_even_length():
    # n = 4, p = [2^60, 1, -2^60, 1]: s = 0 + (2^60 + 1) = 2^60;
    # s + (-2^60 + 1) = 2^60 - 2^60 = 0.  Left-to-right or odd-style: 1.
    x = np.array([2.0**60, 1.0, -(2.0**60), 1.0])
    assert oracle.ddot(4, x, 1, np.ones(4), 1) == 0.0


def test_dot_strided_and_empty():
    x = np.arange(1.0, 13.0)
    assert oracle.ddot(0, x, 1, x, 1) == 0.0
    # x[0], x[3], x[6] dot y[0], y[2], y[4] with y = x: 1*1 + 4*3 + 7*5 = 48
    assert oracle.ddot(3, x, 3, x, 2) == 48.0


# ---------------------------------------------------------------- LU
@pytest.mark.parametrize("m,n,lda", LU_CASES)
def test_lu_matches_openblas(m, n, lda):
    a = rand_buf(42, m, n, lda)
    a0 = a.copy()
    piv, sign = oracle.dgetrf(m, n, a, lda)
    lu_ref, ipiv, info = lapack.dgetrf(view(a0, m, n, lda))
    np.testing.assert_allclose(view(a, m, n, lda), lu_ref, rtol=0, atol=1e-12)
    perm = ipiv_to_perm(ipiv, m)
    assert np.array_equal(piv, perm)
    nswaps = int(np.sum(ipiv != np.arange(len(ipiv))))
    assert sign == (-1) ** nswaps
    # padding rows [m, lda) never touched
    for j in range(n):
        assert np.array_equal(a[j * lda + m:(j + 1) * lda], a0[j * lda + m:(j + 1) * lda])


def test_lu_zero_matrix():
    a = np.zeros(9)
    piv, sign = oracle.dgetrf(3, 3, a, 3)
    assert np.array_equal(piv, [0, 1, 2]) and sign == 1 and np.all(a == 0)


def test_lu_singular_1_to_9():
    a = np.arange(1.0, 10.0)
    a0 = a.copy()
    piv, sign = oracle.dgetrf(3, 3, a, 3)
    lu_ref, ipiv, _ = lapack.dgetrf(view(a0, 3, 3, 3))
    np.testing.assert_allclose(view(a, 3, 3, 3), lu_ref, atol=1e-12)
    assert np.array_equal(piv, ipiv_to_perm(ipiv, 3))


@pytest.mark.parametrize("n", [32, 64, 100])
def test_lu_residual_and_bounds(n):
    a = rand_buf(7 + n, n, n, n, dist="normal")
    a0 = view(a, n, n, n)
    piv, sign = oracle.dgetrf(n, n, a, n)
    lu = view(a, n, n, n)
    assert sorted(piv.tolist()) == list(range(n))
    assert np.all(np.abs(np.tril(lu, -1)) <= 1.0)
    assert lu_residual_ratio(a0, lu, piv) <= 10.0
    det_sign = np.sign(np.linalg.det(a0))
    assert sign * np.prod(np.sign(np.diag(lu))) == det_sign


def _exact_lu_pivots(mat):
    """Partial-pivot Doolittle LU in exact rationals; first maximum wins."""
    m, n = len(mat), len(mat[0])
    a = [[Fraction(v) for v in row] for row in mat]
    perm = list(range(m))
    for j in range(min(m, n)):
        p = max(range(j, m), key=lambda i: (abs(a[i][j]), -i))
        if p != j:
            a[p], a[j] = a[j], a[p]
            perm[p], perm[j] = perm[j], perm[p]
        if a[j][j] != 0:
            for i in range(j + 1, m):
                f = a[i][j] / a[j][j]
                a[i][j] = f
                for c in range(j + 1, n):
                    a[i][c] -= f * a[j][c]
    return perm, a


@pytest.mark.parametrize("seed", range(40))
def test_lu_brute_force_exact_rationals(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 6))
    mat = rng.integers(-9, 10, size=(n, n)).astype(float)
    perm_exact, lu_exact = _exact_lu_pivots(mat.tolist())
    a = mat.T.reshape(-1).copy()
    piv, _ = oracle.dgetrf(n, n, a, n)
    # Ties in exact arithmetic can be broken by rounding in floating point;
    # only compare when every pivot choice was strict in exact arithmetic.
    lu = view(a, n, n, n)
    np.testing.assert_allclose(lu, np.array(lu_exact, dtype=float), rtol=1e-12, atol=1e-12)
    assert list(piv) == perm_exact


def test_lu_pivot_rule_nan_inf_ties():
    # column [NaN, 5, 3]: nothing compares greater than |NaN| -> p = 0
    a = np.array([np.nan, 5.0, 3.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0])
    piv, _ = oracle.dgetrf(3, 3, a, 3)
    assert piv[0] == 0
    # [1, NaN, 3] -> NaN never chosen -> p = 2
    a = np.array([1.0, np.nan, 3.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0])
    piv, _ = oracle.dgetrf(3, 3, a, 3)
    assert piv[0] == 2
    # [1, inf, -inf] -> first infinite magnitude wins -> p = 1
    a = np.array([1.0, np.inf, -np.inf, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0])
    piv, _ = oracle.dgetrf(3, 3, a, 3)
    assert piv[0] == 1
    # exact tie [1, -2, 2] -> lowest index -> p = 1
    a = np.array([1.0, -2.0, 2.0, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0])
    piv, _ = oracle.dgetrf(3, 3, a, 3)
    assert piv[0] == 1


def test_lu_permutation_matrix_exact():
    p = np.array([2, 0, 3, 1])
    mat = np.eye(4)[p]
    a = mat.T.reshape(-1).copy()
    piv, sign = oracle.dgetrf(4, 4, a, 4)
    lu = view(a, 4, 4, 4)
    assert np.array_equal(lu, np.eye(4))
    assert np.array_equal(mat[piv], np.eye(4))
    assert sign == round(np.linalg.det(mat))


# ---------------------------------------------------------------- Cholesky
@pytest.mark.parametrize("n,lda", [(1, 1), (1, 3), (3, 3), (3, 5), (11, 11), (11, 17), (16, 16),
                                   (16, 24), (23, 23), (23, 31), (33, 40), (100, 104)])
def test_cholesky_matches_openblas(n, lda):
    buf = spd_upper_buf(100 + n, n, lda)
    full = view(buf, n, n, lda)
    a_sym = np.triu(full) + np.triu(full, 1).T
    info = oracle.dpotrf(n, buf, lda)
    assert info == 0
    l = view(buf, n, n, lda)
    c, pinfo = lapack.dpotrf(a_sym, lower=1)
    assert pinfo == 0
    np.testing.assert_allclose(np.tril(l), np.tril(c), rtol=0, atol=1e-12 * max(1.0, np.abs(c).max()))
    assert np.all(np.triu(l, 1) == 0.0)  # strict upper zeroed exactly
    assert chol_residual_ratio(a_sym, np.tril(l)) <= 10.0
    for j in range(n):  # padding untouched (NaN stays NaN)
        assert np.all(np.isnan(buf[j * lda + n:(j + 1) * lda]))


@pytest.mark.parametrize("seed", range(12))
def test_cholesky_maybe_singular_info(seed):
    # CholeskyTests.cs:139-176/400-416: random (usually indefinite) upper triangle
    n = 2 + seed
    rng = np.random.default_rng(seed)
    up = np.triu(2 * rng.random((n, n)) - 1) + np.diag(rng.random(n) * (seed % 3))
    a_sym = np.triu(up) + np.triu(up, 1).T
    buf = np.full(n * n, np.nan)
    for j in range(n):
        buf[j * n:j * n + j + 1] = up[: j + 1, j]
    info = oracle.dpotrf(n, buf, n)
    _, pinfo = lapack.dpotrf(a_sym, lower=1)
    assert info == pinfo


def test_cholesky_closed_form_2x2():
    # [[4, 2], [2, 5]] -> L = [[2, 0], [1, 2]] exactly
    buf = np.array([4.0, np.nan, 2.0, 5.0])
    assert oracle.dpotrf(2, buf, 2) == 0
    assert np.array_equal(buf, [2.0, 1.0, 0.0, 2.0])


def test_cholesky_nan_diagonal_fails():
    buf = np.array([np.nan, np.nan, 1.0, 1.0])
    assert oracle.dpotrf(2, buf, 2) == 1


# ---------------------------------------------------------------- TRSV
@pytest.mark.parametrize("uplo", [oracle.UPPER, oracle.LOWER])
@pytest.mark.parametrize("trans", [oracle.NOTRANS, oracle.TRANS])
@pytest.mark.parametrize("n,lda,incx", [(1, 1, 1), (3, 3, 2), (7, 9, 1), (16, 16, 3), (33, 40, 1), (100, 100, 2)])
def test_trsv_matches_openblas(uplo, trans, n, lda, incx):
    a = rand_buf(42, n, n, lda)
    # make the triangle well conditioned and NON-diagonal
    for j in range(n):
        a[j * lda + j] = 4.0 + abs(a[j * lda + j]) * n ** 0.5
    x = rand_buf(57, 1, n * incx, 1)
    x0 = x.copy()
    assert oracle.dtrsv(uplo, trans, n, a, lda, x, incx) == 0
    tri = view(a, n, n, lda)
    tri = np.triu(tri) if uplo == oracle.UPPER else np.tril(tri)
    op = tri.T if trans == oracle.TRANS else tri
    expect = np.linalg.solve(op, x0[::incx][:n])
    np.testing.assert_allclose(x[::incx][:n], expect, rtol=0, atol=1e-11)
    # elements between strides untouched
    mask = np.ones(x.size, bool)
    mask[::incx] = False
    assert np.array_equal(x[mask], x0[mask])


def test_getrs_composition_solves():
    n = 50
    a = rand_buf(3, n, n, n, dist="normal")
    a0 = view(a, n, n, n)
    piv, _ = oracle.dgetrf(n, n, a, n)
    b = np.random.default_rng(4).standard_normal(n)
    x = oracle.dgetrs(n, a, n, piv, b)
    np.testing.assert_allclose(a0 @ x, b, atol=1e-10)


# ---------------------------------------------------------------- GEMM
GEMM_CASES = [(1, 1, 1, 1, 1, 1), (3, 4, 5, 3, 5, 3), (3, 4, 5, 7, 9, 8), (8, 16, 4, 8, 16, 8),
              (17, 9, 13, 20, 20, 19), (32, 32, 32, 32, 32, 32)]


def _op(mat, t):
    return {111: mat, 112: mat.T, 113: mat.conj().T, 114: mat.conj()}[t]


@pytest.mark.parametrize("ta", [111, 112])
@pytest.mark.parametrize("tb", [111, 112])
@pytest.mark.parametrize("m,n,k,lda0,ldb0,ldc", GEMM_CASES)
def test_dgemm_matches_numpy(ta, tb, m, n, k, lda0, ldb0, ldc):
    ra, ca = (m, k) if ta == 111 else (k, m)
    rb, cb = (k, n) if tb == 111 else (n, k)
    lda, ldb, ldc = max(lda0, ra), max(ldb0, rb), max(ldc, m)
    a = rand_buf(42, ra, ca, lda)
    b = rand_buf(57, rb, cb, ldb)
    c = rand_buf(0, m, n, ldc)
    c0 = c.copy()
    assert oracle.dgemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc) == 0
    expect = _op(view(a, ra, ca, lda), ta) @ _op(view(b, rb, cb, ldb), tb)
    np.testing.assert_allclose(view(c, m, n, ldc), expect, rtol=0, atol=1e-12)
    for j in range(n):
        assert np.array_equal(c[j * ldc + m:(j + 1) * ldc], c0[j * ldc + m:(j + 1) * ldc])
    # alpha/beta extension
    c2 = c0.copy()
    oracle.dgemm(ta, tb, m, n, k, -0.75, a, lda, b, ldb, 0.5, c2, ldc)
    np.testing.assert_allclose(view(c2, m, n, ldc), -0.75 * expect + 0.5 * view(c0, m, n, ldc), atol=1e-12)


# ---------------------------------------------------------------- LU inverse
# (n, lda) of LuInverseTests.cs:75-84, plus larger orders
GETRI_CASES = [(1, 1), (1, 3), (2, 2), (2, 3), (3, 3), (3, 5), (5, 5), (5, 8), (10, 10), (10, 17), (33, 40)]


def _lu_of(seed, n, lda):
    a = rand_buf(seed, n, n, lda)
    a0 = a.copy()
    piv, _ = oracle.dgetrf(n, n, a, lda)
    return a0, a, piv


@pytest.mark.parametrize("n,lda", GETRI_CASES)
def test_dgetri_times_a_is_identity(n, lda):
    """LuInverseTests.cs:85-136: A * inv(A) = I within 1e-12, padding rows
    untouched; also against LAPACK's inverse."""
    a0, a, piv = _lu_of(42, n, lda)
    assert oracle.dgetri(n, a, lda, piv) == 0
    inv = view(a, n, n, lda)
    np.testing.assert_allclose(view(a0, n, n, lda) @ inv, np.eye(n), rtol=0, atol=1e-12)
    np.testing.assert_allclose(inv, np.linalg.inv(view(a0, n, n, lda)), rtol=1e-10, atol=1e-12)
    for j in range(n):
        assert np.array_equal(a[j * lda + n:(j + 1) * lda], a0[j * lda + n:(j + 1) * lda], equal_nan=True)


@pytest.mark.parametrize("n,lda", [(1, 1), (5, 8), (10, 17), (33, 40), (64, 64)])
def test_dgetri_column_is_lu_solve_of_unit_vector(n, lda):
    """The identity the batched inverse rests on: column j of LuInverse's
    result is, bit for bit, the LU-solve composition (LuInverse.cs:169-192
    for one right-hand side) applied to e_j."""
    _, a, piv = _lu_of(7, n, lda)
    lu = a.copy()
    assert oracle.dgetri(n, a, lda, piv) == 0
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        x = oracle.dgetrs(n, lu, lda, piv, e)
        assert np.array_equal(x.view(np.uint64), a[j * lda:j * lda + n].view(np.uint64)), j


def test_dgetri_permutation_and_zero_diagonal():
    """A permutation matrix factors exactly (L = I, U = I), so its inverse is
    its transpose exactly; a zero diagonal entry of U returns j + 1 and
    leaves the input unchanged (the reference throws before writing)."""
    n = 6
    perm = np.array([3, 0, 5, 1, 2, 4])
    p = np.zeros((n, n))
    p[np.arange(n), perm] = 1.0
    a = p.T.reshape(-1).copy()  # column-major
    piv, _ = oracle.dgetrf(n, n, a, n)
    assert oracle.dgetri(n, a, n, piv) == 0
    assert np.array_equal(view(a, n, n, n), p.T)
    _, a, piv = _lu_of(3, 5, 5)
    a[2 * 5 + 2] = 0.0
    before = a.copy()
    assert oracle.dgetri(5, a, 5, piv) == 3
    assert np.array_equal(a, before)


# ---------------------------------------------------------------- GEMV
# (m, n, lda, incx, incy) of BlasTests.MulMatVecDouble.cs:9-22 and :56-63
GEMV_CASES = [(1, 1, 1, 1, 1), (1, 1, 3, 2, 4), (2, 2, 4, 3, 3), (3, 3, 5, 6, 7), (1, 5, 3, 2, 4),
              (4, 1, 6, 5, 7), (4, 9, 5, 3, 2), (8, 3, 9, 2, 2), (1, 5, 2, 2, 2), (4, 1, 5, 3, 4),
              (70, 129, 71, 2, 1)]


@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("m,n,lda,incx,incy", GEMV_CASES)
def test_dgemv_matches_openblas(trans, m, n, lda, incx, incy):
    """Blas.MulMatVec against OpenBLAS dgemv at 1e-12, as the reference's
    tests check it; elements of y between strides stay untouched."""
    from scipy.linalg import blas
    a = rand_buf(42, m, n, lda)
    xl, yl = (m, n) if trans == 112 else (n, m)
    rng = np.random.default_rng(57)
    x = rng.standard_normal(xl * incx)
    y = rng.standard_normal(yl * incy)
    y0 = y.copy()
    assert oracle.dgemv(trans, m, n, a, lda, x, incx, y, incy) == 0
    want = blas.dgemv(1.0, view(a, m, n, lda), x[::incx], trans=int(trans == 112))
    np.testing.assert_allclose(y[::incy], want, rtol=0, atol=1e-12)
    mask = np.ones(y.size, dtype=bool)
    mask[::incy] = False
    assert np.array_equal(y[mask], y0[mask])


def test_dgemv_summation_order():
    """Every output is one Internals.Dot: NoTrans over a row (stride lda),
    Trans over a column.  Row [2^60, 1, -2^60, 1] (even length) gives 0 in
    MatFlat's pairing and 1 left to right; column [1, 2^60, -2^60, 1, 1]
    (odd length) gives 3, where left to right gives 2."""
    big = 2.0**60
    lda = 2
    a = np.array([big, np.nan, 1.0, np.nan, -big, np.nan, 1.0, np.nan])  # 1 x 4, lda 2
    y = np.zeros(1)
    assert oracle.dgemv(111, 1, 4, a, lda, np.ones(4), 1, y, 1) == 0
    assert y[0] == 0.0
    col = np.array([1.0, big, -big, 1.0, 1.0])
    assert oracle.dgemv(112, 5, 1, col, 5, np.ones(5), 1, y, 1) == 0
    assert y[0] == 3.0
    # A^T x with A = [[1, 2], [3, 4]] (column-major [1, 3, 2, 4]), x = [1, 1]
    y2 = np.zeros(2)
    oracle.dgemv(112, 2, 2, np.array([1.0, 3.0, 2.0, 4.0]), 2, np.ones(2), 1, y2, 1)
    assert y2.tolist() == [4.0, 6.0]
    oracle.dgemv(111, 2, 2, np.array([1.0, 3.0, 2.0, 4.0]), 2, np.ones(2), 1, y2, 1)
    assert y2.tolist() == [3.0, 7.0]
    assert oracle.dgemv(113, 2, 2, np.ones(4), 2, np.ones(2), 1, y2, 1) == -1


@pytest.mark.parametrize("ta", [111, 112, 113, 114])
@pytest.mark.parametrize("tb", [111, 112, 113, 114])
@pytest.mark.parametrize("m,n,k,lda0,ldb0,ldc", GEMM_CASES[:5])
def test_zgemm_matches_numpy(ta, tb, m, n, k, lda0, ldb0, ldc):
    ra, ca = (m, k) if ta in (111, 114) else (k, m)
    rb, cb = (k, n) if tb in (111, 114) else (n, k)
    lda, ldb, ldc = max(lda0, ra), max(ldb0, rb), max(ldc, m)
    a = rand_cbuf(42, ra, ca, lda)
    b = rand_cbuf(57, rb, cb, ldb)
    c = rand_cbuf(0, m, n, ldc)
    c0 = c.copy()
    assert oracle.zgemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc) == 0
    expect = _op(view(a, ra, ca, lda), ta) @ _op(view(b, rb, cb, ldb), tb)
    np.testing.assert_allclose(view(c, m, n, ldc), expect, rtol=0, atol=1e-12)
    for j in range(n):
        assert np.array_equal(c[j * ldc + m:(j + 1) * ldc], c0[j * ldc + m:(j + 1) * ldc])
    al, be = 0.3 - 0.8j, -0.6 + 0.1j
    c2 = c0.copy()
    oracle.zgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, c2, ldc)
    np.testing.assert_allclose(view(c2, m, n, ldc), al * expect + be * view(c0, m, n, ldc), atol=1e-12)


def test_zgemm_sequential_order():
    # Dot_NN sums sequentially from Complex.Zero (Blas.MulMatMat.cs:598-607):
    # 1, 2^60, -2^60 -> ((0+1)+2^60)-2^60 = 0; a pairwise sum would give 1.
    a = np.array([1.0, 2.0**60, -(2.0**60)], dtype=np.complex128)  # 1 x 3, lda 1
    b = np.ones(3, dtype=np.complex128)  # 3 x 1
    c = np.zeros(1, dtype=np.complex128)
    oracle.zgemm(111, 111, 1, 1, 3, 1.0, a, 1, b, 3, 0.0, c, 1)
    assert c[0] == 0.0


def test_zgemm_conj_forms_closed_form():
    # (1+2i) * (3+4i) forms: NN -5+10i, CN conj(x)y = 11-2i, NC x conj(y) = 11+2i, CC -5-10i
    x = np.array([1 + 2j])
    y = np.array([3 + 4j])
    for ta, tb, want in [(111, 111, -5 + 10j), (114, 111, 11 - 2j), (111, 114, 11 + 2j), (114, 114, -5 - 10j)]:
        c = np.zeros(1, dtype=np.complex128)
        oracle.zgemm(ta, tb, 1, 1, 1, 1.0, x, 1, y, 1, 0.0, c, 1)
        assert c[0] == want


# ---------------------------------------------------------------- posv and adjoint
def test_posv_solves():
    n = 40
    buf = spd_upper_buf(5, n, n)
    full = view(buf, n, n, n)
    a_sym = np.triu(full) + np.triu(full, 1).T
    b = np.random.default_rng(6).standard_normal(n)
    x = b.copy()
    assert oracle.dposv(n, buf, n, x) == 0
    np.testing.assert_allclose(x, np.linalg.solve(a_sym, b), rtol=1e-12, atol=1e-13)


def test_posv_backward_finite_differences():
    n = 6
    rng = np.random.default_rng(9)
    bm = rng.standard_normal((n, n))
    a_sym = bm @ bm.T + n * np.eye(n)
    b = rng.standard_normal(n)
    w = rng.standard_normal(n)

    def loss(a_upper_sym, bb):
        buf = np.full(n * n, np.nan)
        for j in range(n):
            buf[j * n:j * n + j + 1] = a_upper_sym[: j + 1, j]
        x = bb.copy()
        assert oracle.dposv(n, buf, n, x) == 0
        return w @ x, buf, x

    _, lbuf, x = loss(a_sym, b)
    abar, bbar = oracle.dposv_backward(n, lbuf, n, x, w)
    abar = view(abar, n, n, n)
    h = 1e-6
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd = (loss(a_sym, b + e)[0] - loss(a_sym, b - e)[0]) / (2 * h)
        assert abs(fd - bbar[i]) <= 1e-8 * max(1.0, abs(bbar[i]))
        for j in range(i, n):
            s = np.zeros((n, n))
            s[i, j] = s[j, i] = h
            fd = (loss(a_sym + s, b)[0] - loss(a_sym - s, b)[0]) / (2 * h)
            want = abar[i, j] if i == j else abar[i, j] + abar[j, i]
            assert abs(fd - want) <= 1e-8 * max(1.0, abs(want))
    # closed form: bbar = A^-1 w, Abar = -sym(bbar x^T)
    np.testing.assert_allclose(bbar, np.linalg.solve(a_sym, w), rtol=1e-12)
    np.testing.assert_allclose(abar, -0.5 * (np.outer(bbar, x) + np.outer(x, bbar)), rtol=1e-12)


def test_batch_drivers_match_single():
    ns = np.array([3, 10, 17, 8], dtype=np.int32)
    ldas = ((ns + 7) // 8 * 8).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(ldas.astype(np.int64) * ns)[:-1]])
    a = np.concatenate([spd_upper_buf(i, int(n), int(l)) for i, (n, l) in enumerate(zip(ns, ldas))])
    xoffs = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]])
    x = np.random.default_rng(1).standard_normal(int(ns.sum()))
    a1, x1 = a.copy(), x.copy()
    info = oracle.dposv_batch(ns, ldas, offs, a, xoffs, x, threads=2)
    assert np.all(info == 0)
    for i in range(len(ns)):
        o, l, nn = int(offs[i]), int(ldas[i]), int(ns[i])
        seg = a1[o:o + l * nn].copy()
        xs = x1[xoffs[i]:xoffs[i] + nn].copy()
        oracle.dposv(nn, seg, l, xs)
        assert np.array_equal(np.nan_to_num(seg, nan=7.0), np.nan_to_num(a[o:o + l * nn], nan=7.0))
        assert np.array_equal(xs, x[xoffs[i]:xoffs[i] + nn])


def test_lu_solve_batch_matches_single_and_solves():
    """The threaded LU-solve driver gives the same bits as per-matrix
    dgetrf + dgetrs, and x solves A x = b to a small backward error."""
    count, n = 6, 13
    rng = np.random.default_rng(11)
    a = rng.standard_normal(count * n * n)
    b = rng.standard_normal(count * n)
    a0 = a.copy()
    piv, x = oracle.dgesv_batch(count, n, n, n * n, a, b, threads=3)
    for i in range(count):
        seg = a0[i * n * n:(i + 1) * n * n].copy()
        p, _ = oracle.dgetrf(n, n, seg, n)
        xs = oracle.dgetrs(n, seg, n, p, b[i * n:(i + 1) * n])
        assert np.array_equal(seg, a[i * n * n:(i + 1) * n * n])
        assert np.array_equal(p, piv[i * n:(i + 1) * n])
        assert np.array_equal(xs, x[i * n:(i + 1) * n])
        am = a0[i * n * n:(i + 1) * n * n].reshape(n, n).T
        r = b[i * n:(i + 1) * n] - am @ xs
        assert np.abs(r).max() <= 1e-13 * (np.abs(am).sum(1).max() * np.abs(xs).max() + 1)


def test_gemm_and_backward_batch_drivers_match_single():
    """The threaded GEMM and adjoint drivers give the same bits as their
    per-problem calls."""
    rng = np.random.default_rng(12)
    for cplx in (0, 1):
        w = 2 if cplx else 1
        cnt = 5
        m, n, k = (rng.integers(1, 9, size=cnt) for _ in range(3))
        ops = [111, 112, 113] if cplx else [111, 112]
        ta, tb = rng.choice(ops, size=cnt), rng.choice(ops, size=cnt)
        ra, ca = np.where(ta == 111, m, k), np.where(ta == 111, k, m)
        rb, cb = np.where(tb == 111, k, n), np.where(tb == 111, n, k)
        lda, ldb, ldc = ra + 1, rb + 2, m + 1
        sa, sb, sc = lda * ca * w, ldb * cb * w, ldc * n * w
        oa, ob, oc = (np.concatenate([[0], np.cumsum(s)[:-1]]) for s in (sa, sb, sc))
        a, b, c = (rng.standard_normal(int(s.sum())) for s in (sa, sb, sc))
        al, be = rng.uniform(-1, 1, cnt * w), rng.uniform(-1, 1, cnt * w)
        c0 = c.copy()
        info = oracle.gemm_vbatch(cplx, ta, tb, m, n, k, al, a, oa, lda, b, ob, ldb, be, c, oc, ldc, threads=2)
        assert np.all(info == 0)
        for i in range(cnt):
            ci = c0[oc[i]:oc[i] + sc[i]].copy()
            ai, bi = a[oa[i]:oa[i] + sa[i]], b[ob[i]:ob[i] + sb[i]]
            if cplx:
                cz = ci.view(np.complex128)
                oracle.zgemm(int(ta[i]), int(tb[i]), int(m[i]), int(n[i]), int(k[i]), complex(al[2 * i], al[2 * i + 1]),
                             ai.view(np.complex128), int(lda[i]), bi.view(np.complex128), int(ldb[i]),
                             complex(be[2 * i], be[2 * i + 1]), cz, int(ldc[i]))
            else:
                oracle.dgemm(int(ta[i]), int(tb[i]), int(m[i]), int(n[i]), int(k[i]), al[i], ai.copy(), int(lda[i]),
                             bi.copy(), int(ldb[i]), be[i], ci, int(ldc[i]))
            assert np.array_equal(ci, c[oc[i]:oc[i] + sc[i]]), (cplx, i)
    cnt, n, lda = 4, 7, 8
    ls = np.concatenate([spd_upper_buf(40 + i, n, lda) for i in range(cnt)])
    for i in range(cnt):
        seg = ls[i * n * lda:(i + 1) * n * lda]
        assert oracle.dpotrf(n, seg, lda) == 0
    x, wv = rng.standard_normal(cnt * n), rng.standard_normal(cnt * n)
    abar, bbar = oracle.dposv_backward_batch(cnt, n, lda, n * lda, ls, x, wv, threads=2)
    for i in range(cnt):
        ab, bb = oracle.dposv_backward(n, ls[i * n * lda:(i + 1) * n * lda].copy(), lda, x[i * n:(i + 1) * n],
                                       wv[i * n:(i + 1) * n], ldab=lda)
        assert np.array_equal(ab, abar[i * n * lda:(i + 1) * n * lda])
        assert np.array_equal(bb, bbar[i * n:(i + 1) * n])
