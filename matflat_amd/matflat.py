"""MatFlat's hot-path signatures, executed on the GPU.

``Factorization.Lu``, ``Factorization.Cholesky``, ``Blas.SolveTriangular``
and ``Blas.MulMatMat`` keep the reference's argument lists, argument checks
(same order, same conditions) and exceptions, with device tensors in place
of pointers: ``a`` is a 1-D float64 or complex128 CUDA tensor holding a
column-major matrix with leading dimension ``lda``, exactly the buffer the
reference's pointer addresses; a complex128 tensor selects the reference's
Complex overload.  Each call runs the batched
HIP kernel with a batch of one and waits for it (the reference is
synchronous).

Reference signatures:
* Factorization.Lu(int m, int n, double* / Complex* a, int lda, int* piv) -> int
  (MatFlat/Factorization.Lu.cs:122-161, :199-238)
* Factorization.Cholesky(int n, double* / Complex* a, int lda)
  (MatFlat/Factorization.Cholesky.cs:114-168, :195-249)
* Factorization.LuInverse(int n, double* a, int lda, int* piv)
  (MatFlat/Factorization.LuInverse.cs:105-192)
* Factorization.Gevd(int n, double* a, int lda, double* b, int ldb, double* w)
  (MatFlat/Factorization.Gevd.cs:145-186, GevdCore :330-376)
* Blas.SolveTriangular(Uplo, Transpose, int n, double* / Complex* a, int lda, x, int incx)
  (MatFlat/Blas.SolveTriangular.cs:162-242, :280-395)
* Blas.MulMatMat(Transpose, Transpose, int m, int n, int k, double* a, int lda,
  double* b, int ldb, double* c, int ldc)  (MatFlat/Blas.MulMatMat.cs:226-368, :407-575)
* Blas.MulMatVec(Transpose, int m, int n, double* a, int lda, double* x, int incx,
  double* y, int incy)  (MatFlat/Blas.MulMatVec.cs:138-206)
"""
from __future__ import annotations

from enum import IntEnum

import torch

from . import hip
from .hip import ArgumentError, MatFlatException, MatrixFactorizationException  # noqa: F401


class Transpose(IntEnum):
    """Transpose.cs:8-29 (CBLAS values)."""

    NoTrans = 111
    Trans = 112
    ConjTrans = 113
    ConjNoTrans = 114


class Uplo(IntEnum):
    """Uplo.cs:8-19 (CBLAS values)."""

    Upper = 121
    Lower = 122


class ArgumentException(ArgumentError):
    def __init__(self, message: str, param: str | None = None):
        ValueError.__init__(self, f"{message} (Parameter '{param}')" if param else message)
        self.param = param
        self.position = 0


class ArgumentNullException(ArgumentException):
    def __init__(self, param: str):
        super().__init__("Value cannot be null.", param)


def _dev(t: torch.Tensor) -> torch.device:
    return t.device


_exact_handle: hip.Handle | None = None


def _exact() -> hip.Handle:
    """The handle every drop-in call runs on: its own, always in exact math
    mode, so a caller that switches the default handle to FAST mode does not
    change these results (bit-identical to MatFlat, partial writes on a
    failed Cholesky as MatFlat leaves them before its throw at
    Factorization.Cholesky.cs:156)."""
    global _exact_handle
    if _exact_handle is None:
        _exact_handle = hip.Handle()
        _exact_handle.set_math_mode(hip.MATH_EXACT)
    _exact_handle.bind_stream()
    return _exact_handle


def _ptrs(*ts: torch.Tensor) -> torch.Tensor:
    return torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device=ts[0].device)


def _i32(v, device) -> torch.Tensor:
    return torch.tensor(v if isinstance(v, (list, tuple)) else [v], dtype=torch.int32, device=device)


def _need(buf: torch.Tensor | None, elems: int, name: str) -> None:
    if buf is not None and buf.numel() < elems:
        raise ArgumentException(f"The buffer is too small ({buf.numel()} < {elems}).", name)


class Factorization:
    @staticmethod
    def Lu(m: int, n: int, a: torch.Tensor | None, lda: int, piv: torch.Tensor | None) -> int:
        if m <= 0:
            raise ArgumentException("The number of rows must be greater than or equal to one.", "m")
        if n <= 0:
            raise ArgumentException("The number of columns must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < m:
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "lda")
        if piv is None:
            raise ArgumentNullException("piv")
        _need(a, lda * (n - 1) + m, "a")
        _need(piv, m, "piv")
        dev = _dev(a)
        sign = torch.zeros(1, dtype=torch.int32, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        if a.is_complex():  # Factorization.Lu.cs:199-238
            hip.zgetrf_vbatched(_i32(m, dev), _i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(piv), sign, info, m, n,
                                handle=_exact())
        else:
            hip.dgetrf_strided_batched(1, m, n, a, lda, lda * n, piv, m, sign, info, handle=_exact())
        torch.cuda.synchronize()
        return int(sign.item())

    @staticmethod
    def Cholesky(n: int, a: torch.Tensor | None, lda: int) -> None:
        if n <= 0:
            raise ArgumentException("The order of the matrix must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < n:
            raise ArgumentException(
                "The leading dimension must be greater than or equal to the order of the matrix.", "lda")
        _need(a, lda * (n - 1) + n, "a")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        if a.is_complex():  # Factorization.Cholesky.cs:195-249
            hip.zpotrf_vbatched(_i32(n, dev), _ptrs(a), _i32(lda, dev), info, n, handle=_exact())
        else:
            hip.dpotrf_vbatched(_i32(n, dev), _ptrs(a), _i32(lda, dev), info, n, handle=_exact())
        if int(info.item()) != 0:
            raise MatrixFactorizationException(
                "Cholesky decomposition failed. The matrix must be positive definite.")

    @staticmethod
    def LuInverse(n: int, a: torch.Tensor | None, lda: int, piv: torch.Tensor | None) -> None:
        """A^-1 from Lu's factors and pivots, in place
        (Factorization.LuInverse.cs:105-192, double)."""
        if n <= 0:
            raise ArgumentException("The order of the matrix must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < n:
            raise ArgumentException(
                "The leading dimension must be greater than or equal to the order of the matrix.", "lda")
        if piv is None:
            raise ArgumentNullException("piv")
        _need(a, lda * (n - 1) + n, "a")
        _need(piv, n, "piv")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        hip.dgetri_vbatched(_i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(piv), info, n, handle=_exact())
        if int(info.item()) != 0:
            raise MatrixFactorizationException("The matrix is not invertible.")


    @staticmethod
    def Gevd(n: int, a: torch.Tensor | None, lda: int, b: torch.Tensor | None, ldb: int,
             w: torch.Tensor | None) -> None:
        """A v = w B v for symmetric A and symmetric positive definite B (upper
        triangles read): a -> eigenvectors, b -> L of B = L L^T, w -> the
        eigenvalues, descending (Factorization.Gevd.cs:145-186, double)."""
        if n <= 0:
            raise ArgumentException("The order of the matrix must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < n:
            raise ArgumentException(
                "The leading dimension must be greater than or equal to the order of the matrix.", "lda")
        if b is None:
            raise ArgumentNullException("b")
        if ldb < n:
            raise ArgumentException(
                "The leading dimension must be greater than or equal to the order of the matrix.", "ldb")
        if w is None:
            raise ArgumentNullException("w")
        _need(a, lda * (n - 1) + n, "a")
        _need(b, ldb * (n - 1) + n, "b")
        _need(w, n, "w")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        hip.dgevd_vbatched(_i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(b), _i32(ldb, dev), _ptrs(w), info, n,
                           handle=_exact())
        code = int(info.item())
        if code == n + 1:  # Factorization.Svd.cs:990-993
            raise MatrixFactorizationException("SVD failed. The solution did not converge.")
        if code != 0:  # Factorization.Gevd.cs:336-343
            raise MatFlatException("The right-hand side matrix is not positive definite.")

class Blas:
    @staticmethod
    def SolveTriangular(uplo: int, transa: int, n: int, a: torch.Tensor | None, lda: int, x: torch.Tensor | None,
                        incx: int) -> None:
        if n <= 0:
            raise ArgumentException("The order of the matrix must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < n:
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "lda")
        if x is None:
            raise ArgumentNullException("x")
        if incx <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "incx")
        cplx = a.is_complex()
        # the reference checks transa inside the uplo branch, so an invalid
        # uplo is reported first (Blas.SolveTriangular.cs:190-241, :306-394)
        if uplo not in (Uplo.Upper, Uplo.Lower):
            raise ArgumentException("Invalid enum value.", "uplo")
        valid = (Transpose.NoTrans, Transpose.Trans, Transpose.ConjTrans, Transpose.ConjNoTrans) if cplx else \
            (Transpose.NoTrans, Transpose.Trans)
        if transa not in valid:
            raise ArgumentException("Invalid enum value.", "transa")
        _need(a, lda * (n - 1) + n, "a")
        _need(x, incx * (n - 1) + 1, "x")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        if cplx:  # Blas.SolveTriangular.cs:280-395
            hip.ztrsv_vbatched(int(uplo), int(transa), _i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(x),
                               _i32(incx, dev), info, n, handle=_exact())
        else:
            hip.dtrsv_vbatched(int(uplo), int(transa), hip.NONUNIT, _i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(x),
                               _i32(incx, dev), info, n, handle=_exact())
        torch.cuda.synchronize()

    @staticmethod
    def MulMatMat(transa: int, transb: int, m: int, n: int, k: int, a: torch.Tensor | None, lda: int,
                  b: torch.Tensor | None, ldb: int, c: torch.Tensor | None, ldc: int) -> None:
        cplx = a is not None and a.is_complex()
        valid = (111, 112, 113, 114) if cplx else (111, 112)
        if m <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "m")
        if n <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "n")
        if k <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "k")
        if a is None:
            raise ArgumentNullException("a")
        if transa not in valid:
            raise ArgumentException("Invalid enum value.", "transa")
        if lda < (k if transa in (112, 113) else m):
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "lda")
        if b is None:
            raise ArgumentNullException("b")
        if transb not in valid:
            raise ArgumentException("Invalid enum value.", "transb")
        if ldb < (n if transb in (112, 113) else k):
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "ldb")
        if c is None:
            raise ArgumentNullException("c")
        if ldc < m:
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "ldc")
        if a.data_ptr() == c.data_ptr() or b.data_ptr() == c.data_ptr():
            raise ArgumentException("In-place operation is not supported.")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        one = torch.tensor([1.0, 0.0] if cplx else [1.0], dtype=torch.float64, device=dev)
        zero = torch.zeros(2 if cplx else 1, dtype=torch.float64, device=dev)
        hip.gemm_vbatched(_i32(int(transa), dev), _i32(int(transb), dev), _i32(m, dev), _i32(n, dev), _i32(k, dev),
                          one, _ptrs(a), _i32(lda, dev), _ptrs(b), _i32(ldb, dev), zero, _ptrs(c), _i32(ldc, dev),
                          info, m, n, k, complex_=cplx, handle=_exact())
        torch.cuda.synchronize()

    @staticmethod
    def MulMatVec(trans: int, m: int, n: int, a: torch.Tensor | None, lda: int, x: torch.Tensor | None, incx: int,
                  y: torch.Tensor | None, incy: int) -> None:
        """y = op(A) x, Blas.MulMatVec.cs:138-206 (double): checks in the
        reference's order, the transpose last."""
        if m <= 0:
            raise ArgumentException("The number of rows must be greater than or equal to one.", "m")
        if n <= 0:
            raise ArgumentException("The number of columns must be greater than or equal to one.", "n")
        if a is None:
            raise ArgumentNullException("a")
        if lda < m:
            raise ArgumentException("The leading dimension must be greater than or equal to the number of rows.",
                                    "lda")
        if x is None:
            raise ArgumentNullException("x")
        if incx <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "incx")
        if y is None:
            raise ArgumentNullException("y")
        if incy <= 0:
            raise ArgumentException("The value must be greater than or equal to one.", "incy")
        if a.data_ptr() == y.data_ptr() or x.data_ptr() == y.data_ptr():
            raise ArgumentException("In-place operation is not supported.")
        if trans not in (111, 112):
            raise ArgumentException("Invalid enum value.", "trans")
        xl, yl = (m, n) if trans == 112 else (n, m)
        _need(a, lda * (n - 1) + m, "a")
        _need(x, incx * (xl - 1) + 1, "x")
        _need(y, incy * (yl - 1) + 1, "y")
        dev = _dev(a)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        hip.dgemv_vbatched(_i32(int(trans), dev), _i32(m, dev), _i32(n, dev), _ptrs(a), _i32(lda, dev), _ptrs(x),
                           _i32(incx, dev), _ptrs(y), _i32(incy, dev), info, m, n, handle=_exact())
        torch.cuda.synchronize()
