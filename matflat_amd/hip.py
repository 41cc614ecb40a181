"""Batched device API over libmatflat_hip.so (the MatFlat.Hip surface).

Every function here takes torch tensors that already live on the GPU and
calls the C ABI on torch's current stream; nothing is computed on the CPU.
Per-matrix arrays (sizes, leading dimensions, pointer arrays, info) are
device int32/int64 tensors; ``VBatch`` builds them from a packed buffer and
offsets.

Errors: a negative status from the library raises ``ArgumentError`` (the
analogue of MatFlat's ArgumentException, e.g. Factorization.Lu.cs:124-147)
with the library's message; other failures raise ``HipError``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import (BLOCKED_MAX_N, CONJNOTRANS, CONJTRANS, EXACT_MAX_N, LOWER, MATH_EXACT, MATH_FAST,  # noqa: F401
                   NONUNIT, NOTRANS, SMALL_MAX_N, TRANS, UNIT, UPPER)


class MatFlatException(Exception):
    """MatFlatException.cs:9-30."""


class MatrixFactorizationException(MatFlatException):
    """MatrixFactorizationException.cs:9-30: the factorization failed."""


class ArgumentError(ValueError):
    """An invalid argument (MatFlat raises ArgumentException/ArgumentNullException)."""

    def __init__(self, position: int, message: str):
        super().__init__(f"argument {position}: {message}")
        self.position = position


class HipError(RuntimeError):
    pass


def _check(status: int) -> None:
    if status == _lib.MFH_OK:
        return
    msg = _lib.last_error()
    if -1000 < status < 0:
        raise ArgumentError(-status, msg)
    raise HipError(f"libmatflat_hip status {status}: {msg}")


def _p(t) -> int:
    if t is None:
        return 0
    if isinstance(t, int):
        return t
    return t.data_ptr()


class Handle:
    """Owns an mfh_handle bound to torch's current stream at call time.

    With ``tuned`` (the default) the launch shapes of the committed tuning
    table for this GPU's architecture (matflat_amd/tuning/<arch>.json,
    written by ``python -m matflat_amd.autotune``) are applied."""

    def __init__(self, device: int | None = None, tuned: bool = True):
        self.lib = _lib.load()
        if device is not None:
            torch.cuda.set_device(device)
        torch.cuda.init()
        h = ctypes.c_void_p()
        _check(self.lib.mfh_create(ctypes.byref(h)))
        self.h = h
        if tuned:
            from . import autotune
            arch = torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
            table = autotune.load(arch)
            if table is not None:
                autotune.apply(self, table)

    def set_chol_shape(self, bucket: int, lanes: int, rows_per_lane: int) -> None:
        """Exact Cholesky: lanes per matrix and rows per lane for a 16-row
        bucket ((0, 0) = built-in)."""
        _check(self.lib.mfh_set_chol_shape(self.h, int(bucket), int(lanes), int(rows_per_lane)))

    def bind_stream(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(self.lib.mfh_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)))

    def set_math_mode(self, mode: int) -> None:
        """MATH_EXACT (bit-identical to MatFlat) or MATH_FAST (FMA, blocked)."""
        _check(self.lib.mfh_set_math_mode(self.h, int(mode)))

    def math_mode(self) -> int:
        m = ctypes.c_int()
        _check(self.lib.mfh_get_math_mode(self.h, ctypes.byref(m)))
        return m.value

    def set_fast_waves(self, row_blocks: int, waves: int) -> None:
        """FAST Cholesky: wavefronts per matrix (1, 2, 4; 0 = built-in) for up to 8*row_blocks rows."""
        _check(self.lib.mfh_set_fast_waves(self.h, int(row_blocks), int(waves)))

    def set_bucket_streams(self, streams: int) -> None:
        """FAST bucket launches forked over this many streams (1 = in order)."""
        _check(self.lib.mfh_set_bucket_streams(self.h, int(streams)))

    def bucket_streams(self) -> int:
        s = ctypes.c_int()
        _check(self.lib.mfh_get_bucket_streams(self.h, ctypes.byref(s)))
        return s.value

    def fast_waves(self, row_blocks: int) -> int:
        w = ctypes.c_int()
        _check(self.lib.mfh_get_fast_waves(self.h, int(row_blocks), ctypes.byref(w)))
        return w.value

    def set_timing(self, enable: bool) -> None:
        """Bracket every kernel launch on this handle with HIP events."""
        _check(self.lib.mfh_set_timing(self.h, 1 if enable else 0))

    def timing(self) -> list[tuple[int, int, float]]:
        """Waits for and returns the timed launches as (kind, rows, ms), then clears them."""
        cap = 4096
        n = ctypes.c_int()
        kind = (ctypes.c_int * cap)()
        rows = (ctypes.c_int * cap)()
        ms = (ctypes.c_float * cap)()
        _check(self.lib.mfh_get_timing(self.h, cap, ctypes.byref(n), kind, rows, ms))
        return [(kind[i], rows[i], ms[i]) for i in range(min(cap, n.value))]

    def synchronize(self) -> None:
        _check(self.lib.mfh_synchronize(self.h))

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.mfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default: Handle | None = None


def default_handle() -> Handle:
    global _default
    if _default is None:
        _default = Handle()
    _default.bind_stream()
    return _default


@dataclass
class VBatch:
    """Device-side per-matrix arrays for a batch packed in one buffer.

    ``buf`` is a 1-D device tensor; matrix i starts at element ``offsets[i]``.
    ``ptrs`` is the device int64 array of element addresses the C ABI takes.
    """

    buf: torch.Tensor
    offsets: torch.Tensor  # int64, device
    ptrs: torch.Tensor     # int64, device

    @staticmethod
    def of(buf: torch.Tensor, offsets) -> "VBatch":
        offs = torch.as_tensor(offsets, dtype=torch.int64, device=buf.device)
        ptrs = buf.data_ptr() + offs * buf.element_size()
        return VBatch(buf, offs, ptrs.contiguous())


def _i32(x, device) -> torch.Tensor:
    return torch.as_tensor(x, dtype=torch.int32, device=device).contiguous()


# ---------------------------------------------------------------------------
def dgetrf_vbatched(m, n, a_ptrs, lda, piv_ptrs, sign, info, max_m, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgetrf_vbatched(h.h, int(m.numel()), _p(m), _p(n), _p(a_ptrs), _p(lda), _p(piv_ptrs),
                                     _p(sign), _p(info), int(max_m), int(max_n)))


def dgetrf_strided_batched(count, m, n, a, lda, stride_a, piv, stride_piv, sign, info, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgetrf_strided_batched(h.h, int(count), int(m), int(n), _p(a), int(lda), int(stride_a),
                                            _p(piv), int(stride_piv), _p(sign), _p(info)))


def dpotrf_vbatched(n, a_ptrs, lda, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dpotrf_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(info), int(max_n)))


def dposv_vbatched(n, a_ptrs, lda, x_ptrs, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dposv_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(x_ptrs), _p(info),
                                    int(max_n)))


def dtrsv_vbatched(uplo, trans, diag, n, a_ptrs, lda, x_ptrs, incx, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dtrsv_vbatched(h.h, int(uplo), int(trans), int(diag), int(n.numel()), _p(n), _p(a_ptrs),
                                    _p(lda), _p(x_ptrs), _p(incx), _p(info), int(max_n)))


def dgetrs_vbatched(n, lu_ptrs, lda, piv_ptrs, b_ptrs, x_ptrs, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgetrs_vbatched(h.h, int(n.numel()), _p(n), _p(lu_ptrs), _p(lda), _p(piv_ptrs), _p(b_ptrs),
                                     _p(x_ptrs), _p(info), int(max_n)))


def gemm_vbatched(transa, transb, m, n, k, alpha, a_ptrs, lda, b_ptrs, ldb, beta, c_ptrs, ldc, info,
                  max_m, max_n, max_k, complex_: bool = False, handle: Handle | None = None):
    h = handle or default_handle()
    fn = h.lib.mfh_zgemm_vbatched if complex_ else h.lib.mfh_dgemm_vbatched
    _check(fn(h.h, int(m.numel()), _p(transa), _p(transb), _p(m), _p(n), _p(k), _p(alpha), _p(a_ptrs), _p(lda),
              _p(b_ptrs), _p(ldb), _p(beta), _p(c_ptrs), _p(ldc), _p(info), int(max_m), int(max_n), int(max_k)))


def dgetri_vbatched(n, a_ptrs, lda, piv_ptrs, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgetri_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(piv_ptrs), _p(info),
                                     int(max_n)))


def dgemv_vbatched(trans, m, n, a_ptrs, lda, x_ptrs, incx, y_ptrs, incy, info, max_m, max_n,
                   handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgemv_vbatched(h.h, int(m.numel()), _p(trans), _p(m), _p(n), _p(a_ptrs), _p(lda), _p(x_ptrs),
                                    _p(incx), _p(y_ptrs), _p(incy), _p(info), int(max_m), int(max_n)))


def zgetrf_vbatched(m, n, a_ptrs, lda, piv_ptrs, sign, info, max_m, max_n, handle: Handle | None = None):
    """Complex128 LU (mfh_zgetrf_vbatched); a_ptrs address interleaved (re, im) matrices."""
    h = handle or default_handle()
    _check(h.lib.mfh_zgetrf_vbatched(h.h, int(m.numel()), _p(m), _p(n), _p(a_ptrs), _p(lda), _p(piv_ptrs),
                                     _p(sign), _p(info), int(max_m), int(max_n)))


def zpotrf_vbatched(n, a_ptrs, lda, info, max_n, handle: Handle | None = None):
    """Complex128 Cholesky A = L L^H (mfh_zpotrf_vbatched)."""
    h = handle or default_handle()
    _check(h.lib.mfh_zpotrf_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(info), int(max_n)))


def ztrsv_vbatched(uplo, trans, n, a_ptrs, lda, x_ptrs, incx, info, max_n, handle: Handle | None = None):
    """Complex128 triangular solve, every transpose (mfh_ztrsv_vbatched)."""
    h = handle or default_handle()
    _check(h.lib.mfh_ztrsv_vbatched(h.h, int(uplo), int(trans), int(n.numel()), _p(n), _p(a_ptrs), _p(lda),
                                    _p(x_ptrs), _p(incx), _p(info), int(max_n)))


def dgesv_mixed_vbatched(n, a_ptrs, lda, b_ptrs, x_ptrs, max_iter, tol_per_n, iters, berr, info, max_n,
                         handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgesv_mixed_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(b_ptrs), _p(x_ptrs),
                                          int(max_iter), float(tol_per_n), _p(iters), _p(berr), _p(info),
                                          int(max_n)))


def dgesv_vbatched(n, a_ptrs, lda, b_ptrs, x_ptrs, berr, info, max_n, handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dgesv_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(b_ptrs), _p(x_ptrs),
                                    _p(berr), _p(info), int(max_n)))


def dposv_vbatched_noalloc(n, a_ptrs, lda, x_ptrs, info, max_n, handle: Handle | None = None):
    """The SPD solve of dposv_vbatched on the library's static scratch: no
    device allocation (mfh_dposv_vbatched_noalloc; count <= NOALLOC_MAX_COUNT)."""
    h = handle or default_handle()
    _check(h.lib.mfh_dposv_vbatched_noalloc(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(x_ptrs), _p(info),
                                            int(max_n)))


def dgetrf_blocked_vbatched(n, a_ptrs, lda, piv_ptrs, info, max_n, handle: Handle | None = None):
    """FAST-mode LU with partial pivoting for n <= 512 (mfh_dgetrf_blocked_vbatched)."""
    h = handle or default_handle()
    _check(h.lib.mfh_dgetrf_blocked_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(piv_ptrs), _p(info),
                                             int(max_n)))


def dposv_backward_vbatched(n, l_ptrs, lda, x_ptrs, xbar_ptrs, abar_ptrs, ldabar, bbar_ptrs, info, max_n,
                            handle: Handle | None = None):
    h = handle or default_handle()
    _check(h.lib.mfh_dposv_backward_vbatched(h.h, int(n.numel()), _p(n), _p(l_ptrs), _p(lda), _p(x_ptrs),
                                             _p(xbar_ptrs), _p(abar_ptrs), _p(ldabar), _p(bbar_ptrs), _p(info),
                                             int(max_n)))


def dpotrf_backward_vbatched(n, l_ptrs, ldl, lbar_ptrs, ldlbar, abar_ptrs, ldabar, info, max_n,
                             handle: Handle | None = None):
    """Adjoint of the batched Cholesky: abar = dloss/dA (full symmetric) from
    L and lbar = dloss/dL."""
    h = handle or default_handle()
    _check(h.lib.mfh_dpotrf_backward_vbatched(h.h, int(n.numel()), _p(n), _p(l_ptrs), _p(ldl), _p(lbar_ptrs),
                                              _p(ldlbar), _p(abar_ptrs), _p(ldabar), _p(info), int(max_n)))


def dgevd_vbatched(n, a_ptrs, lda, b_ptrs, ldb, w_ptrs, info, max_n, handle: Handle | None = None):
    """Factorization.Gevd (double) over a batch: a -> eigenvectors, b -> L,
    w -> eigenvalues (descending), in MatFlat's operation order."""
    h = handle or default_handle()
    _check(h.lib.mfh_dgevd_vbatched(h.h, int(n.numel()), _p(n), _p(a_ptrs), _p(lda), _p(b_ptrs), _p(ldb),
                                    _p(w_ptrs), _p(info), int(max_n)))


class Plan:
    """One batched call captured into a hipGraph (mfh_plan_create_*) and
    replayed by execute() on the handle's stream.  The device arrays the call
    was given must outlive the plan."""

    def __init__(self, create, args, handle: Handle | None = None):
        self.handle = handle or default_handle()
        self._keep = args
        p = ctypes.c_void_p()
        _check(create(self.handle.h, *[_p(a) if not isinstance(a, (int, float)) else a for a in args],
                      ctypes.byref(p)))
        self.p = p

    def op(self) -> int:
        v = ctypes.c_int()
        _check(self.handle.lib.mfh_plan_get_op(self.p, ctypes.byref(v)))
        return v.value

    def execute(self) -> None:
        self.handle.bind_stream()
        _check(self.handle.lib.mfh_plan_execute(self.p))

    def close(self) -> None:
        if getattr(self, "p", None) is not None and self.p.value:
            self.handle.lib.mfh_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GemmPlan(Plan):
    """A batched GEMM (gemm_vbatched's arguments) captured once into a
    hipGraph and replayed."""

    def __init__(self, transa, transb, m, n, k, alpha, a_ptrs, lda, b_ptrs, ldb, beta, c_ptrs, ldc, info,
                 max_m, max_n, max_k, complex_: bool = False, handle: Handle | None = None):
        h = handle or default_handle()
        fn = h.lib.mfh_plan_create_zgemm if complex_ else h.lib.mfh_plan_create_dgemm
        super().__init__(fn, (int(m.numel()), transa, transb, m, n, k, alpha, a_ptrs, lda, b_ptrs, ldb, beta, c_ptrs,
                              ldc, info, int(max_m), int(max_n), int(max_k)), h)


class DposvPlan(Plan):
    """A batched SPD solve captured once into a hipGraph and replayed."""

    def __init__(self, n, a_ptrs, lda, x_ptrs, info, max_n, handle: Handle | None = None):
        h = handle or default_handle()
        super().__init__(h.lib.mfh_plan_create_dposv, (int(n.numel()), n, a_ptrs, lda, x_ptrs, info, int(max_n)), h)


def plan_dgetrf(m, n, a_ptrs, lda, piv_ptrs, sign, info, max_m, max_n, handle: Handle | None = None) -> Plan:
    h = handle or default_handle()
    return Plan(h.lib.mfh_plan_create_dgetrf,
                (int(m.numel()), m, n, a_ptrs, lda, piv_ptrs, sign, info, int(max_m), int(max_n)), h)


def plan_dpotrf(n, a_ptrs, lda, info, max_n, handle: Handle | None = None) -> Plan:
    h = handle or default_handle()
    return Plan(h.lib.mfh_plan_create_dpotrf, (int(n.numel()), n, a_ptrs, lda, info, int(max_n)), h)


def plan_dtrsv(uplo, trans, diag, n, a_ptrs, lda, x_ptrs, incx, info, max_n, handle: Handle | None = None) -> Plan:
    h = handle or default_handle()
    return Plan(h.lib.mfh_plan_create_dtrsv,
                (int(uplo), int(trans), int(diag), int(n.numel()), n, a_ptrs, lda, x_ptrs, incx, info, int(max_n)), h)


def plan_dgesv(n, a_ptrs, lda, b_ptrs, x_ptrs, berr, info, max_n, handle: Handle | None = None) -> Plan:
    h = handle or default_handle()
    return Plan(h.lib.mfh_plan_create_dgesv, (int(n.numel()), n, a_ptrs, lda, b_ptrs, x_ptrs, berr, info, int(max_n)),
                h)


def plan_dgesv_mixed(n, a_ptrs, lda, b_ptrs, x_ptrs, max_iter, tol_per_n, iters, berr, info, max_n,
                     handle: Handle | None = None) -> Plan:
    h = handle or default_handle()
    return Plan(h.lib.mfh_plan_create_dgesv_mixed,
                (int(n.numel()), n, a_ptrs, lda, b_ptrs, x_ptrs, int(max_iter), float(tol_per_n), iters, berr, info,
                 int(max_n)), h)
