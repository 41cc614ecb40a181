"""Pins the oracle's generalized eigenproblem (oracle/matflat_oracle.c:
mfo_dnorm, mfo_drotg, mfo_dsvd_u, mfo_dgevd) before the GPU kernel is
checked against it bit for bit.

* Norm: hand-derived sums whose value depends on the pairing, and numpy's
  2-norm to rounding.
* Drotg: OpenBLAS drotg (scipy) for c and s, r = +-hypot(a, b) with the sign
  of the larger operand, and z by the reference BLAS rule.
* SVD (left vectors only): numpy's singular values, U orthogonal,
  U^T A A^T U = diag(s^2), s descending; closed forms: 1x1, and a diagonal
  matrix, whose vectors the algorithm's steps fix exactly (Householder
  vectors (2, 0, ...) give U = diag(-1, ..., -1, 1), then the ordering
  passes permute columns by descending |d|).
* GEVD: the reference test's check A v = w B v (GevdTests.cs:124-156,
  tolerance 1e-9) on its shape list, scipy's eigh(A, B) eigenvalues,
  V^T B V = I, L L^T = B, untouched padding rows and unread lower
  triangles (NaN), the 1x1 and 2x2 closed forms, diagonal A with B = I
  (signs and order fixed as for the SVD), and a B that is not positive
  definite.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.blas as blas

import oracle
from helpers import EPS, chol_residual_ratio, gevd_pair_bufs, view

GEVD_CASES = [(1, 1, 1), (1, 3, 2), (2, 2, 2), (2, 4, 3), (3, 3, 3), (3, 4, 5), (5, 5, 5), (5, 7, 8),
              (10, 10, 10), (10, 17, 11)]  # GevdTests.cs:94-105


def bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


# ------------------------------------------------------------ Norm / Drotg
def test_norm_pairing():
    # odd length: x0^2 first, then (x1^2 + x2^2); even: 0 + (x0^2 + x1^2) + ...
    x = np.array([1e8, 1.0, 1.0])
    assert oracle.dnorm(3, x) == np.sqrt(1e16 + (1.0 + 1.0))
    y = np.array([1.0, 1e-8, 1e-8, 1.0])
    s = 0.0 + (1.0 + 1e-16)
    s = s + (1e-16 + 1.0)
    assert oracle.dnorm(4, y) == np.sqrt(s)
    assert oracle.dnorm(2, np.array([3.0, 4.0])) == 5.0
    assert oracle.dnorm(0, np.zeros(1)) == 0.0


@pytest.mark.parametrize("n", [1, 2, 7, 64, 129])
def test_norm_numpy(n):
    x = np.random.default_rng(n).standard_normal(n)
    assert abs(oracle.dnorm(n, x) - np.linalg.norm(x)) <= 4 * n * EPS * np.linalg.norm(x)


@pytest.mark.parametrize("a,b", [(3.0, 4.0), (-4.0, 3.0), (4.0, -3.0), (-3.0, -4.0), (1e-300, 2e-300), (0.5, 0.0),
                                 (0.0, -2.0), (1.0, 1.0), (-1.0, 1.0), (7.25, -1e-3)])
def test_drotg(a, b):
    r, z, c, s = oracle.drotg(a, b)
    cr, sr = blas.drotg(a, b)
    assert abs(c - cr) <= 2 * EPS and abs(s - sr) <= 2 * EPS
    roe = a if abs(a) > abs(b) else b
    assert abs(abs(r) - np.hypot(a, b)) <= 2 * EPS * np.hypot(a, b)
    assert np.sign(r) == np.sign(roe)
    assert c * a + s * b == pytest.approx(r, rel=4 * EPS)
    if abs(a) > abs(b):
        assert z == s
    elif c != 0.0:
        assert z == 1.0 / c
    else:
        assert z == 1.0


def test_drotg_zero():
    assert oracle.drotg(0.0, 0.0) == (0.0, 0.0, 1.0, 0.0)


# ------------------------------------------------------------ SVD
def svd_u(a_mat):
    n = a_mat.shape[0]
    a = np.asfortranarray(a_mat).reshape(-1, order="F").copy()
    u = np.full(n * n, np.nan)
    info, s = oracle.dsvd_u(n, a, n, u, n)
    return info, s, u.reshape(n, n, order="F")


@pytest.mark.parametrize("n", [2, 3, 5, 8, 17, 40, 64])
def test_svd_numpy(n):
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal((n, n))
    info, s, u = svd_u(a)
    assert info == 0
    ref = np.linalg.svd(a, compute_uv=False)
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs(s - ref)) <= 16 * n * EPS * ref[0]
    assert np.max(np.abs(u.T @ u - np.eye(n))) <= 16 * n * EPS
    g = u.T @ a @ a.T @ u
    assert np.max(np.abs(g - np.diag(s * s))) <= 64 * n * EPS * ref[0] ** 2


def test_svd_1x1():
    for v in (2.5, -2.5, 0.0):
        info, s, u = svd_u(np.array([[v]]))
        assert info == 0 and s[0] == abs(v) and u[0, 0] == 1.0


@pytest.mark.parametrize("d", [[3.0, -1.0, 2.0, 0.5], [1.0, 2.0], [-5.0, 4.0, 3.0], [0.25, 8.0, -0.5, 2.0, 1.0]])
def test_svd_diagonal_closed_form(d):
    d = np.array(d)
    n = d.size
    info, s, u = svd_u(np.diag(d))
    assert info == 0
    order = sorted(range(n), key=lambda i: -abs(d[i]))  # distinct |d|: the bubble order
    assert np.array_equal(s, np.abs(d[order]))
    want = np.zeros((n, n))
    for col, i in enumerate(order):
        want[i, col] = 1.0 if i == n - 1 else -1.0
    assert np.array_equal(u, want)


def test_svd_rank_deficient():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((9, 4))
    a = x @ x.T  # rank 4
    info, s, u = svd_u(a)
    ref = np.linalg.svd(a, compute_uv=False)
    assert info == 0 and np.max(np.abs(s - ref)) <= 200 * EPS * ref[0]
    assert np.max(np.abs(u.T @ u - np.eye(9))) <= 200 * EPS


# ------------------------------------------------------------ GEVD
def run_gevd(a_buf, b_buf, n, lda, ldb):
    a, b = a_buf.copy(), b_buf.copy()
    w = np.zeros(n)
    info = oracle.dgevd(n, a, lda, b, ldb, w)
    return info, a, b, w


@pytest.mark.parametrize("n,lda,ldb", GEVD_CASES + [(16, 19, 16), (33, 33, 40), (64, 64, 64)])
def test_gevd_reference_cases(n, lda, ldb):
    a_buf, b_buf, am, bm = gevd_pair_bufs(42 + n, n, lda, ldb, shift=0.5)
    info, a, b, w = run_gevd(a_buf, b_buf, n, lda, ldb)
    assert info == 0
    v = view(a, n, n, lda)
    # the reference test's own check (GevdTests.cs:124-156): A v_j = w_j B v_j within 1e-9
    for j in range(n):
        assert np.max(np.abs(am @ v[:, j] - w[j] * (bm @ v[:, j]))) <= 1e-9 * max(1.0, np.abs(am).max())
    ref = sl.eigh(am, bm, eigvals_only=True)[::-1]
    assert np.max(np.abs(w - ref)) <= 1e-12 * n * np.abs(ref).max()
    assert np.max(np.abs(v.T @ bm @ v - np.eye(n))) <= 1e-10
    l = np.tril(view(b, n, n, ldb))
    assert np.array_equal(view(b, n, n, ldb), l)  # strict upper zeroed
    assert chol_residual_ratio(bm, l) <= 10
    # padding rows untouched (GevdTests.cs:158-176)
    for buf, out, ld in ((a_buf, a, lda), (b_buf, b, ldb)):
        pad = np.ones(ld * n, bool)
        for j in range(n):
            pad[j * ld : j * ld + n] = False
        assert np.array_equal(bits(buf[pad]), bits(out[pad]))
    assert np.all(np.isfinite(v)) and np.all(np.isfinite(w))


def test_gevd_1x1_closed_form():
    a = np.array([3.0, 9.0])
    b = np.array([7.0])
    w = np.zeros(1)
    assert oracle.dgevd(1, a, 2, b, 1, w) == 0
    l = np.sqrt(7.0)
    assert b[0] == l and w[0] == (3.0 / l) / l and a[0] == 1.0 / l and a[1] == 9.0


def test_gevd_2x2_closed_form():
    am = np.array([[4.0, 1.0], [1.0, 3.0]])
    bm = np.array([[2.0, 0.5], [0.5, 1.0]])
    # det(A - w B) = 0: (4-2w)(3-w) - (1-0.5w)^2 = 1.75 w^2 - 9 w + 11
    qa, qb, qc = 1.75, -9.0, 11.0
    r = np.sqrt(qb * qb - 4 * qa * qc)
    ref = np.array([(-qb + r) / (2 * qa), (-qb - r) / (2 * qa)])
    a = np.array([4.0, np.nan, 1.0, 3.0])
    b = np.array([2.0, np.nan, 0.5, 1.0])
    w = np.zeros(2)
    assert oracle.dgevd(2, a, 2, b, 2, w) == 0
    assert np.max(np.abs(w - ref)) <= 8 * EPS * ref[0]


def test_gevd_diagonal_closed_form():
    d = np.array([3.0, 1.0, 5.0, 0.5, 2.0])
    n = d.size
    a = np.diag(d).reshape(-1, order="F").copy()
    b = np.eye(n).reshape(-1, order="F").copy()
    w = np.zeros(n)
    assert oracle.dgevd(n, a, n, b, n, w) == 0
    order = sorted(range(n), key=lambda i: -d[i])
    assert np.array_equal(w, d[order])
    want = np.zeros((n, n))
    for col, i in enumerate(order):
        want[i, col] = 1.0 if i == n - 1 else -1.0
    assert np.array_equal(a.reshape(n, n, order="F"), want)


def test_gevd_b_not_positive_definite():
    n, lda, ldb = 6, 7, 6
    a_buf, b_buf, am, bm = gevd_pair_bufs(5, n, lda, ldb, shift=0.5)
    b_buf = b_buf.copy()
    b_buf[3 + 3 * ldb] = -50.0  # row 3's pivot goes negative
    info, a, b, w = run_gevd(a_buf, b_buf, n, lda, ldb)
    want_b = b_buf.copy()
    assert oracle.dpotrf(n, want_b, ldb) == 4 == info
    assert np.array_equal(bits(b), bits(want_b))  # the Cholesky's partial writes
    assert np.array_equal(bits(a), bits(a_buf))  # a untouched
    assert np.all(w == 0)


def test_gevd_planted_transpose_fails():
    """A mistake the checks above would catch: eigenvectors L^-1 U instead of
    L^-T U fail A v = w B v."""
    n = 5
    a_buf, b_buf, am, bm = gevd_pair_bufs(11, n, n, n, shift=0.5)
    info, a, b, w = run_gevd(a_buf, b_buf, n, n, n)
    l = np.tril(view(b, n, n, n))
    u = l.T @ view(a, n, n, n)  # recover U
    wrong = np.linalg.solve(l, u)
    res = max(np.max(np.abs(am @ wrong[:, j] - w[j] * (bm @ wrong[:, j]))) for j in range(n))
    assert res > 1e-6


def test_gevd_nan_does_not_converge():
    """A NaN in A's upper triangle: the iteration never meets its test and
    gives up (the reference throws "SVD failed"); w is not written."""
    n = 8
    a_buf, b_buf, _, _ = gevd_pair_bufs(304, n, n + 1, n + 2, shift=0.25)
    a_buf[2 + 5 * (n + 1)] = np.nan
    info, a, b, w = run_gevd(a_buf, b_buf, n, n + 1, n + 2)
    assert info == n + 1 and np.all(w == 0)
