"""CPU oracle for the MatFlat hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP kernels.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package (``matflat_amd``) never imports it and has no
CPU fallback.

``libmatflat_oracle.so`` is a C restatement (``matflat_oracle.c``) of:

* ``Internals.Dot``           MatFlat/Internals.cs:265-292
* ``Factorization.Lu``        MatFlat/Factorization.Lu.cs:296-350
* ``Factorization.Cholesky``  MatFlat/Factorization.Cholesky.cs:114-168
* ``Blas.SolveTriangular``    MatFlat/Blas.SolveTriangular.cs:162-242, :397-421
* ``Blas.MulMatMat``          MatFlat/Blas.MulMatMat.cs:226-368 (double),
                              :407-648 (Complex)
* LU solve composition        MatFlat/Factorization.LuInverse.cs:169-192
* Complex LU, Cholesky, triangular solve
                              MatFlat/Factorization.Lu.cs:352-406,
                              Factorization.Cholesky.cs:195-249,
                              Blas.SolveTriangular.cs:280-445, with BCL
                              Complex division (Smith's method)
* ``Factorization.Gevd``      MatFlat/Factorization.Gevd.cs:330-376 with
                              SvdCore (Factorization.Svd.cs:786-1256),
                              Internals.Norm and Internals.Drotg

compiled with ``-O2 -ffp-contract=off -fno-fast-math``.  Pinning (what makes
these functions trustworthy, beyond being written from the reference) lives in
``tests/test_oracle.py``: agreement with OpenBLAS (scipy) at the reference
tests' tolerances, factor residual ratios, permutation/triangularity
invariants, closed forms on special matrices and brute-force permutation
search on tiny inputs.

Buffers are 1-D float64 numpy arrays holding a column-major matrix with
leading dimension ``ld`` (element (i, j) at ``[i + j*ld]``), exactly like the
reference tests' arrays (MatFlatTest/Matrix.cs:9-28).  Complex buffers are
numpy complex128 (interleaved re, im).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmatflat_oracle.so")

NOTRANS, TRANS, CONJTRANS, CONJNOTRANS = 111, 112, 113, 114
UPPER, LOWER = 121, 122

_lib = None


def build() -> None:
    """Compile libmatflat_oracle.so with the oracle's Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        i64 = ctypes.c_int64
        _lib.mfo_ddot.restype = d
        _lib.mfo_ddot.argtypes = [i, p, i, p, i]
        _lib.mfo_dgetrf.restype = i
        _lib.mfo_dgetrf.argtypes = [i, i, p, i, p]
        _lib.mfo_dpotrf.restype = i
        _lib.mfo_dpotrf.argtypes = [i, p, i]
        _lib.mfo_dtrsv.restype = i
        _lib.mfo_dtrsv.argtypes = [i, i, i, p, i, p, i]
        _lib.mfo_dtrsv_unit_lower.restype = None
        _lib.mfo_dtrsv_unit_lower.argtypes = [i, p, i, p, i]
        _lib.mfo_dgetrs.restype = None
        _lib.mfo_dgetrs.argtypes = [i, p, i, p, p, p]
        _lib.mfo_dposv.restype = i
        _lib.mfo_dposv.argtypes = [i, p, i, p]
        _lib.mfo_dgemm.restype = i
        _lib.mfo_dgemm.argtypes = [i, i, i, i, i, d, p, i, p, i, d, p, i]
        _lib.mfo_dgetri.restype = i
        _lib.mfo_dgetri.argtypes = [i, p, i, p]
        _lib.mfo_dgemv.restype = i
        _lib.mfo_dgemv.argtypes = [i, i, i, p, i, p, i, p, i]
        _lib.mfo_zgemm.restype = i
        _lib.mfo_zgemm.argtypes = [i, i, i, i, i, p, p, i, p, i, p, p, i]
        _lib.mfo_dposv_batch.restype = None
        _lib.mfo_dposv_batch.argtypes = [i, p, p, p, p, p, p, p, i]
        _lib.mfo_dgetrf_batch.restype = None
        _lib.mfo_dgetrf_batch.argtypes = [i, i, i, i, i64, p, p, p, i]
        _lib.mfo_dgesv_batch.restype = None
        _lib.mfo_dgesv_batch.argtypes = [i, i, i, i64, p, p, p, p, i]
        _lib.mfo_gemm_vbatch.restype = None
        _lib.mfo_gemm_vbatch.argtypes = [i, i] + [p] * 17 + [i]
        _lib.mfo_dposv_backward_batch.restype = None
        _lib.mfo_dposv_backward_batch.argtypes = [i, i, i, i64, p, p, p, p, p, i]
        _lib.mfo_zdiv.restype = None
        _lib.mfo_zdiv.argtypes = [d, d, d, d, p, p]
        _lib.mfo_zdot.restype = None
        _lib.mfo_zdot.argtypes = [i, i, p, i, p, i, p]
        _lib.mfo_zgetrf.restype = i
        _lib.mfo_zgetrf.argtypes = [i, i, p, i, p]
        _lib.mfo_zpotrf.restype = i
        _lib.mfo_zpotrf.argtypes = [i, p, i]
        _lib.mfo_ztrsv.restype = i
        _lib.mfo_ztrsv.argtypes = [i, i, i, p, i, p, i]
        _lib.mfo_dposv_backward.restype = None
        _lib.mfo_dposv_backward.argtypes = [i, p, i, p, p, p, i, p]
        _lib.mfo_dnorm.restype = d
        _lib.mfo_dnorm.argtypes = [i, p]
        _lib.mfo_drotg.restype = None
        _lib.mfo_drotg.argtypes = [p, p, p, p]
        _lib.mfo_dsvd_u.restype = i
        _lib.mfo_dsvd_u.argtypes = [i, p, i, p, p, i, p, p, p]
        _lib.mfo_dgevd.restype = i
        _lib.mfo_dgevd.argtypes = [i, p, i, p, i, p, p, p]
        _lib.mfo_dpotrf_backward.restype = None
        _lib.mfo_dpotrf_backward.argtypes = [i, p, i, p, i, p, i]
        _lib.mfo_dpotrf_backward_batch.restype = None
        _lib.mfo_dpotrf_backward_batch.argtypes = [i, i, i, i64, p, p, p, i]
        _lib.mfo_dgevd_batch.restype = None
        _lib.mfo_dgevd_batch.argtypes = [i, p, p, p, p, p, p, p, p, p, p, i]
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def ddot(n, x, incx, y, incy) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().mfo_ddot(n, _ptr(x), incx, _ptr(y), incy)


def dgetrf(m: int, n: int, a: np.ndarray, lda: int):
    """In-place LU of the (lda*n,) buffer ``a``.  Returns (piv int32[m], sign)."""
    assert a.dtype == np.float64 and a.size >= lda * (n - 1) + m
    piv = np.zeros(m, dtype=np.int32)
    sign = lib().mfo_dgetrf(m, n, _ptr(a), lda, _ptr(piv))
    return piv, sign


def dpotrf(n: int, a: np.ndarray, lda: int) -> int:
    """In-place Cholesky; returns info (0 or j+1 of the failing row)."""
    assert a.dtype == np.float64
    return lib().mfo_dpotrf(n, _ptr(a), lda)


def dtrsv(uplo: int, trans: int, n: int, a: np.ndarray, lda: int, x: np.ndarray, incx: int = 1) -> int:
    return lib().mfo_dtrsv(uplo, trans, n, _ptr(a), lda, _ptr(x), incx)


def dtrsv_unit_lower(n, a, lda, x, incx=1) -> None:
    lib().mfo_dtrsv_unit_lower(n, _ptr(a), lda, _ptr(x), incx)


def dgetrs(n, lu, lda, piv, b) -> np.ndarray:
    x = np.zeros(n, dtype=np.float64)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().mfo_dgetrs(n, _ptr(lu), lda, _ptr(piv), _ptr(b), _ptr(x))
    return x


def dposv(n, a, lda, x) -> int:
    return lib().mfo_dposv(n, _ptr(a), lda, _ptr(x))


def dgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> int:
    return lib().mfo_dgemm(transa, transb, m, n, k, float(alpha), _ptr(a), lda, _ptr(b), ldb,
                           float(beta), _ptr(c), ldc)


def dgetri(n, a, lda, piv) -> int:
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    return lib().mfo_dgetri(n, _ptr(a), lda, _ptr(piv))


def dgemv(trans, m, n, a, lda, x, incx, y, incy) -> int:
    return lib().mfo_dgemv(trans, m, n, _ptr(a), lda, _ptr(x), incx, _ptr(y), incy)


def zgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> int:
    al = np.array([complex(alpha).real, complex(alpha).imag], dtype=np.float64)
    be = np.array([complex(beta).real, complex(beta).imag], dtype=np.float64)
    assert a.dtype == np.complex128 and b.dtype == np.complex128 and c.dtype == np.complex128
    return lib().mfo_zgemm(transa, transb, m, n, k, _ptr(al), _ptr(a), lda, _ptr(b), ldb,
                           _ptr(be), _ptr(c), ldc)


def zdiv(x: complex, y: complex) -> complex:
    """BCL Complex division x / y (Smith's method)."""
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().mfo_zdiv(float(x.real), float(x.imag), float(y.real), float(y.imag), ctypes.byref(re),
                   ctypes.byref(im))
    return complex(re.value, im.value)


def zdot(conj: int, n: int, x, incx: int, y, incy: int) -> complex:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    out = np.zeros(1, dtype=np.complex128)
    lib().mfo_zdot(int(conj), n, _ptr(x), incx, _ptr(y), incy, _ptr(out))
    return complex(out[0])


def zgetrf(m: int, n: int, a: np.ndarray, lda: int):
    """In-place complex LU of the (lda*n,) complex128 buffer ``a``.
    Returns (piv int32[m], sign)."""
    assert a.dtype == np.complex128 and a.size >= lda * (n - 1) + m
    piv = np.zeros(m, dtype=np.int32)
    sign = lib().mfo_zgetrf(m, n, _ptr(a), lda, _ptr(piv))
    return piv, sign


def zpotrf(n: int, a: np.ndarray, lda: int) -> int:
    """In-place complex Cholesky (A = L L^H); returns info (0 or j+1)."""
    assert a.dtype == np.complex128
    return lib().mfo_zpotrf(n, _ptr(a), lda)


def ztrsv(uplo: int, trans: int, n: int, a: np.ndarray, lda: int, x: np.ndarray, incx: int = 1) -> int:
    assert a.dtype == np.complex128 and x.dtype == np.complex128
    return lib().mfo_ztrsv(uplo, trans, n, _ptr(a), lda, _ptr(x), incx)


def dposv_batch(n, lda, off, a, xoff, x, threads=1) -> np.ndarray:
    n = np.ascontiguousarray(n, dtype=np.int32)
    lda = np.ascontiguousarray(lda, dtype=np.int32)
    off = np.ascontiguousarray(off, dtype=np.int64)
    xoff = np.ascontiguousarray(xoff, dtype=np.int64)
    info = np.zeros(n.size, dtype=np.int32)
    lib().mfo_dposv_batch(n.size, _ptr(n), _ptr(lda), _ptr(off), _ptr(a), _ptr(xoff), _ptr(x),
                          _ptr(info), threads)
    return info


def dgetrf_batch(count, m, n, lda, stride, a, threads=1):
    piv = np.zeros(count * m, dtype=np.int32)
    sign = np.zeros(count, dtype=np.int32)
    lib().mfo_dgetrf_batch(count, m, n, lda, stride, _ptr(a), _ptr(piv), _ptr(sign), threads)
    return piv, sign


def dgesv_batch(count, n, lda, stride, a, b, threads=1):
    """LU + LU solve of `count` order-n matrices stride elements apart in a
    (overwritten by the factors); b holds the right-hand sides n apart.
    Returns (piv, x)."""
    piv = np.zeros(count * n, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(count * n, dtype=np.float64)
    lib().mfo_dgesv_batch(count, n, lda, stride, _ptr(a), _ptr(piv), _ptr(b), _ptr(x), threads)
    return piv, x


def gemm_vbatch(cplx, ta, tb, m, n, k, alpha, a, offa, lda, b, offb, ldb, beta, c, offc, ldc, threads=1):
    """mfo_dgemm / mfo_zgemm over a batch; a, b, c are float64 buffers (complex
    entries as (re, im) pairs), offsets count doubles.  Returns info."""
    def i32(v):
        return np.ascontiguousarray(v, dtype=np.int32)

    def i64(v):
        return np.ascontiguousarray(v, dtype=np.int64)

    ta, tb, m, n, k, lda, ldb, ldc = map(i32, (ta, tb, m, n, k, lda, ldb, ldc))
    offa, offb, offc = map(i64, (offa, offb, offc))
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    info = np.zeros(m.size, dtype=np.int32)
    lib().mfo_gemm_vbatch(int(cplx), m.size, _ptr(ta), _ptr(tb), _ptr(m), _ptr(n), _ptr(k), _ptr(alpha), _ptr(a),
                          _ptr(offa), _ptr(lda), _ptr(b), _ptr(offb), _ptr(ldb), _ptr(beta), _ptr(c), _ptr(offc),
                          _ptr(ldc), _ptr(info), threads)
    return info


def dposv_backward_batch(count, n, lda, stride, l, x, w, threads=1):
    """mfo_dposv_backward over `count` order-n factors stride elements apart.
    Returns (abar, bbar) with abar laid out like l."""
    abar = np.zeros(count * stride, dtype=np.float64)
    bbar = np.zeros(count * n, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().mfo_dposv_backward_batch(count, n, lda, stride, _ptr(l), _ptr(x), _ptr(w), _ptr(abar), _ptr(bbar),
                                   threads)
    return abar, bbar


def dposv_backward(n, l, lda, x, w, ldab=None):
    ldab = n if ldab is None else ldab
    abar = np.zeros(ldab * n, dtype=np.float64)
    bbar = np.zeros(n, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().mfo_dposv_backward(n, _ptr(l), lda, _ptr(x), _ptr(w), _ptr(abar), ldab, _ptr(bbar))
    return abar, bbar


def dnorm(n: int, x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return float(lib().mfo_dnorm(n, _ptr(x)))


def drotg(a: float, b: float):
    """Internals.Drotg: returns (r, z, c, s)."""
    v = np.array([a, b, 0.0, 0.0])
    base = v.ctypes.data
    lib().mfo_drotg(base, base + 8, base + 16, base + 24)
    return float(v[0]), float(v[1]), float(v[2]), float(v[3])


def dsvd_u(n: int, a: np.ndarray, lda: int, u: np.ndarray, ldu: int):
    """SvdCore of the square matrix in ``a`` (destroyed) with left vectors
    into ``u``; returns (info, s)."""
    s = np.zeros(n)
    work = np.zeros(3 * n)
    base = work.ctypes.data
    info = lib().mfo_dsvd_u(n, _ptr(a), lda, _ptr(s), _ptr(u), ldu, base, base + 8 * n, base + 16 * n)
    return info, s


def dgevd(n: int, a: np.ndarray, lda: int, b: np.ndarray, ldb: int, w: np.ndarray) -> int:
    """Factorization.Gevd in place: a -> eigenvectors, b -> L, w -> eigenvalues.
    Returns 0, j+1 (B not positive definite at row j) or n+1 (no convergence)."""
    assert a.dtype == np.float64 and b.dtype == np.float64 and w.dtype == np.float64
    c = np.zeros(n * n + 3 * n)
    return lib().mfo_dgevd(n, _ptr(a), lda, _ptr(b), ldb, _ptr(w), _ptr(c), c.ctypes.data + 8 * n * n)


def dgevd_batch(n, lda, ldb, aoff, boff, woff, a, b, w, threads=1) -> np.ndarray:
    n, lda, ldb = (np.ascontiguousarray(v, dtype=np.int32) for v in (n, lda, ldb))
    aoff, boff, woff = (np.ascontiguousarray(v, dtype=np.int64) for v in (aoff, boff, woff))
    info = np.zeros(n.size, dtype=np.int32)
    lib().mfo_dgevd_batch(n.size, _ptr(n), _ptr(lda), _ptr(ldb), _ptr(aoff), _ptr(boff), _ptr(woff), _ptr(a),
                          _ptr(b), _ptr(w), _ptr(info), threads)
    return info


def dpotrf_backward(n: int, l: np.ndarray, ldl: int, lbar: np.ndarray, ldlb: int, ldab: int | None = None) -> np.ndarray:
    """Adjoint of the Cholesky factorization: Abar (full symmetric, ldab x n
    buffer) from L and Lbar = dloss/dL."""
    ldab = ldab or n
    abar = np.zeros(ldab * n)
    l = np.ascontiguousarray(l, dtype=np.float64)
    lbar = np.ascontiguousarray(lbar, dtype=np.float64)
    lib().mfo_dpotrf_backward(n, _ptr(l), ldl, _ptr(lbar), ldlb, _ptr(abar), ldab)
    return abar


def dpotrf_backward_batch(count, n, ld, stride, l, lbar, abar, threads=1) -> None:
    lib().mfo_dpotrf_backward_batch(count, n, ld, stride, _ptr(l), _ptr(lbar), _ptr(abar), threads)
