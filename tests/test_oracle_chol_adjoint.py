"""Pins the oracle's adjoint of the Cholesky factorization
(oracle/matflat_oracle.c: mfo_dpotrf_backward), the checker of
mfh_dpotrf_backward_vbatched.  MatFlat has no derivatives, so the pins are
mathematical:
* central finite differences of loss(A) = sum(Lbar * chol(A)) over symmetric
  perturbations (numpy's Cholesky);
* an independent evaluation of L^-T M L^-1 with numpy's inverse;
* the 1x1 closed form dL/dA = 1 / (2 sqrt(a));
* a planted mistake (Phi without the halved diagonal) fails the differences.
"""
import numpy as np
import pytest

import oracle
from helpers import EPS


def spd(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    return b @ b.T + n * np.eye(n), np.tril(rng.standard_normal((n, n)))


def cm(m, ld=None):
    """column-major buffer with leading dimension ld (padding NaN)."""
    n = m.shape[0]
    ld = ld or n
    buf = np.full(ld * m.shape[1], np.nan)
    for j in range(m.shape[1]):
        buf[j * ld:j * ld + n] = m[:, j]
    return buf


def adjoint(a, lbar, upper_nan=True, ld=None):
    n = a.shape[0]
    l = np.linalg.cholesky(a)
    lb = lbar.copy()
    if upper_nan:
        lb[np.triu_indices(n, 1)] = np.nan  # never read
    ld = ld or n
    out = oracle.dpotrf_backward(n, cm(l, ld), ld, cm(lb, ld), ld, ld)
    return l, out.reshape(n, ld).T[:n, :n]


@pytest.mark.parametrize("n", [2, 5, 12])
def test_finite_differences(n):
    a, lbar = spd(n, n)
    _, g = adjoint(a, lbar)
    h = 1e-6
    rng = np.random.default_rng(100 + n)
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    for idx in rng.choice(len(pairs), size=min(len(pairs), 12), replace=False):
        i, j = pairs[idx]
        e = np.zeros((n, n))
        e[i, j] = e[j, i] = 1.0

        def f(t):
            return np.sum(lbar * np.linalg.cholesky(a + t * e))

        fd = (f(h) - f(-h)) / (2 * h)
        pred = np.sum(g * e)
        assert abs(fd - pred) <= 1e-7 * max(1.0, abs(fd)), (i, j, fd, pred)


@pytest.mark.parametrize("n,ld", [(1, 1), (3, 4), (16, 16), (33, 35), (64, 64)])
def test_independent_formula_and_symmetry(n, ld):
    a, lbar = spd(n, 7 + n)
    l, g = adjoint(a, lbar, ld=ld)
    p = l.T @ lbar
    m = 0.5 * (np.tril(p, -1) + np.tril(p, -1).T) + 0.5 * np.diag(np.diag(p))
    li = np.linalg.inv(l)
    ref = li.T @ m @ li
    assert np.array_equal(g, g.T)
    assert np.max(np.abs(g - ref)) <= 64 * n * EPS * np.abs(ref).max() * np.linalg.cond(l)


def test_one_by_one_closed_form():
    for a, lb in [(4.0, 3.0), (2.0, -1.5)]:
        _, g = adjoint(np.array([[a]]), np.array([[lb]]))
        assert abs(g[0, 0] - lb / (2 * np.sqrt(a))) <= 2 * EPS * abs(g[0, 0])


def test_padding_untouched_and_upper_of_lbar_unread():
    n, ld = 6, 9
    a, lbar = spd(n, 3)
    l = np.linalg.cholesky(a)
    lb = lbar.copy()
    lb[np.triu_indices(n, 1)] = np.nan
    out = np.full(ld * n, -7.0)
    lbuf, lbbuf = cm(l, ld), cm(lb, ld)
    oracle.lib().mfo_dpotrf_backward(n, lbuf.ctypes.data, ld, lbbuf.ctypes.data, ld, out.ctypes.data, ld)
    v = out.reshape(n, ld).T
    assert np.all(v[n:, :] == -7.0) and np.all(np.isfinite(v[:n, :]))


def test_planted_missing_half_diagonal_fails():
    n = 5
    a, lbar = spd(n, 21)
    l = np.linalg.cholesky(a)
    p = l.T @ lbar
    wrong_m = 0.5 * (np.tril(p) + np.tril(p).T)  # diagonal counted fully
    li = np.linalg.inv(l)
    wrong = li.T @ wrong_m @ li
    e = np.zeros((n, n))
    e[2, 2] = 1.0
    h = 1e-6
    fd = (np.sum(lbar * np.linalg.cholesky(a + h * e)) - np.sum(lbar * np.linalg.cholesky(a - h * e))) / (2 * h)
    assert abs(wrong[2, 2] - fd) > 1e-4 * abs(fd)
    _, g = adjoint(a, lbar)
    assert abs(g[2, 2] - fd) <= 1e-7 * max(1.0, abs(fd))
