"""Shared test helpers: seeded inputs, LAPACK pivot conversion, residual ratios.

Inputs follow the reference tests' conventions (MatFlatTest/Matrix.cs:9-28:
a flat column-major buffer of ``n*lda`` entries drawn U(-1, 1), padding
included) but use numpy's PCG64 generator: System.Random's streams are .NET
internals, and the reference tests take their expected values from OpenBLAS
anyway, not from the streams.
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps


def rand_buf(seed: int, rows: int, cols: int, ld: int, dist: str = "uniform") -> np.ndarray:
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        return 2.0 * rng.random(ld * cols) - 1.0
    return rng.standard_normal(ld * cols)


def rand_cbuf(seed: int, rows: int, cols: int, ld: int, dist: str = "uniform") -> np.ndarray:
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        re = 2.0 * rng.random(ld * cols) - 1.0
        im = 2.0 * rng.random(ld * cols) - 1.0
    else:
        re = rng.standard_normal(ld * cols)
        im = rng.standard_normal(ld * cols)
    return re + 1j * im


def view(buf: np.ndarray, rows: int, cols: int, ld: int) -> np.ndarray:
    """rows x cols view (copy) of a column-major buffer."""
    return buf[: ld * cols].reshape(cols, ld).T[:rows, :cols].copy()


def store(buf: np.ndarray, mat: np.ndarray, ld: int) -> None:
    rows, cols = mat.shape
    v = buf[: ld * cols].reshape(cols, ld).T
    v[:rows, :cols] = mat


def ipiv_to_perm(ipiv: np.ndarray, m: int) -> np.ndarray:
    """LAPACK's 0-based swap sequence -> MatFlat's permutation vector
    (piv[i] = original row now at row i; Factorization.Lu.cs:299-337)."""
    perm = np.arange(m, dtype=np.int64)
    for j, p in enumerate(ipiv):
        perm[[j, p]] = perm[[p, j]]
    return perm


def spd_upper_buf(seed: int, n: int, lda: int, shift: float | None = None, nan_lower: bool = True) -> np.ndarray:
    """A = B B^T + shift*I with B ~ N(0,1); upper triangle stored, strict lower
    filled with NaN (CholeskyTests.cs:302-339 fills the unreferenced triangle
    with NaN to prove it is never read)."""
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    a = b @ b.T + (n if shift is None else shift) * np.eye(n)
    buf = np.full(lda * n, np.nan if nan_lower else 0.0)
    full = view(buf, lda, n, lda)
    for j in range(n):
        full[: j + 1, j] = a[: j + 1, j]
    buf[:] = full.T.reshape(-1)
    return buf


def lu_residual_ratio(a0: np.ndarray, lu: np.ndarray, perm: np.ndarray) -> float:
    """||P A - L U||_1 / (n ||A||_1 eps)  (LAPACK dget01)."""
    m, n = a0.shape
    k = min(m, n)
    l = np.tril(lu[:, :k], -1) + np.eye(m, k)
    u = np.triu(lu[:k, :])
    r = a0[perm, :] - l @ u
    an = np.linalg.norm(a0, 1)
    return np.linalg.norm(r, 1) / (max(m, n) * an * EPS) if an > 0 else np.linalg.norm(r, 1)


def chol_residual_ratio(a: np.ndarray, l: np.ndarray) -> float:
    """||L L^T - A||_1 / (n ||A||_1 eps)  (LAPACK dpot01)."""
    n = a.shape[0]
    return np.linalg.norm(l @ l.T - a, 1) / (n * np.linalg.norm(a, 1) * EPS)


def backward_error(a: np.ndarray, x: np.ndarray, b: np.ndarray) -> float:
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    r = b - a @ x
    return np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max())


def hpd_upper_cbuf(seed: int, n: int, lda: int, shift: float | None = None, nan_lower: bool = True) -> np.ndarray:
    """Hermitian positive definite A = B B^H + shift*I with B complex N(0,1);
    upper triangle (diagonal included) stored, strict lower NaN."""
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = b @ b.conj().T + (n if shift is None else shift) * np.eye(n)
    fill = complex(np.nan, np.nan) if nan_lower else 0.0
    buf = np.full(lda * n, fill, dtype=np.complex128)
    for j in range(n):
        buf[j * lda:j * lda + j + 1] = a[: j + 1, j]
    return buf


def gevd_pair_bufs(seed: int, n: int, lda: int, ldb: int, shift: float = 0.0):
    """The pair GevdTests.cs:106-186 builds: A = X X^T and B = Y Y^T (+ shift I)
    from U(-1, 1) matrices, stored in buffers whose padding rows keep their
    random values, with the strict lower triangles set to NaN
    (GevdTests.cs:114-115).  Returns (a_buf, b_buf, A, B)."""
    mats, bufs = [], []
    for s, ld in ((seed, lda), (seed + 1, ldb)):
        buf = rand_buf(s, n, n, ld)
        x = view(buf, n, n, ld)
        m = x @ x.T + (shift * np.eye(n) if s != seed else 0.0)
        store(buf, m, ld)
        v = buf[: ld * n].reshape(n, ld).T
        for j in range(n):
            v[j + 1 : n, j] = np.nan
        mats.append(m)
        bufs.append(buf)
    return bufs[0], bufs[1], mats[0], mats[1]
