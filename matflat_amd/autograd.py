"""torch.autograd bridges for the batched SPD solve and the batched Cholesky
factorization, with their hand-written adjoints.

``spd_solve(a, b)`` solves ``S x = b`` for a batch, where ``S = (A + A^T) / 2`` is
the symmetric part of each ``a[k]`` and must be positive definite.
- The forward pass is the batched Cholesky + two triangular solves, in
  exact mode. It runs on the GPU through libmatflat_hip.so and is
  bit-identical to Factorization.Cholesky followed by
  Blas.SolveTriangular(Lower, NoTrans) and SolveTriangular(Lower, Trans)
  (Factorization.Cholesky.cs:114-168, Blas.SolveTriangular.cs:162-242).
- The backward pass is mfh_dposv_backward_vbatched.

For the backward pass, with ``xbar = dL/dx``:
- ``dL/db = bbar = S^-1 xbar``;
- ``dL/dA = -(bbar x^T + x bbar^T) / 2``. This is the gradient through the
  symmetric part, and the one a symmetric perturbation of A sees.

``cholesky(a)`` returns the batched lower factor L of ``S = (A + A^T) / 2``
(exact mode, bit-identical to Factorization.Cholesky).  Its backward pass is
mfh_dpotrf_backward_vbatched: with ``lbar = dL/dL`` (lower triangle),
``dL/dA = L^-T M L^-1``, ``M = (Phi(L^T lbar) + Phi(L^T lbar)^T) / 2``, Phi the
lower triangle with its diagonal halved -- again the gradient through the
symmetric part.

This module is a test and gradcheck harness; MatFlat itself has no
derivatives.
"""
from __future__ import annotations

import torch

from . import hip


def _check_batch(a: torch.Tensor, b: torch.Tensor) -> None:
    if a.dtype != torch.float64 or b.dtype != torch.float64:
        raise TypeError("spd_solve works in float64")
    if not (a.is_cuda and b.is_cuda):
        raise ValueError("spd_solve needs CUDA (HIP) tensors: there is no CPU path")
    if a.dim() != 3 or a.shape[1] != a.shape[2]:
        raise ValueError("a must have shape (count, n, n)")
    if b.shape != a.shape[:2]:
        raise ValueError("b must have shape (count, n)")
    if a.shape[1] < 1 or a.shape[1] > hip.SMALL_MAX_N:
        raise ValueError(f"n must be in [1, {hip.SMALL_MAX_N}]")


class SpdSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor, b: torch.Tensor):
        _check_batch(a, b)
        count, n = a.shape[0], a.shape[1]
        # S = (A + A^T)/2 is symmetric, so its row-major storage is also its
        # column-major storage (lda = n); the kernel reads one triangle and
        # overwrites the buffer with L.
        l = (0.5 * (a + a.transpose(1, 2))).contiguous()
        x = b.detach().clone().contiguous()
        dev = a.device
        ns = torch.full((count,), n, dtype=torch.int32, device=dev)
        step = torch.arange(count, dtype=torch.int64, device=dev)
        l_ptrs = (l.data_ptr() + 8 * n * n * step).contiguous()
        x_ptrs = (x.data_ptr() + 8 * n * step).contiguous()
        info = torch.zeros(count, dtype=torch.int32, device=dev)
        h = hip.default_handle()
        mode = h.math_mode()
        h.set_math_mode(hip.MATH_EXACT)
        try:
            hip.dposv_vbatched(ns, l_ptrs, ns, x_ptrs, info, n, handle=h)
        finally:
            h.set_math_mode(mode)
        bad = torch.nonzero(info).flatten()
        if bad.numel():
            k = int(bad[0].item())
            raise hip.MatrixFactorizationException(
                f"Cholesky decomposition failed for matrix {k} at row {int(info[k].item()) - 1}. "
                "The matrix must be positive definite.")
        ctx.save_for_backward(l, x)
        return x

    @staticmethod
    def backward(ctx, xbar: torch.Tensor):
        l, x = ctx.saved_tensors
        count, n = l.shape[0], l.shape[1]
        dev = l.device
        xbar = xbar.contiguous()
        abar = torch.empty_like(l)
        bbar = torch.empty_like(x)
        ns = torch.full((count,), n, dtype=torch.int32, device=dev)
        step = torch.arange(count, dtype=torch.int64, device=dev)

        def ptrs(t, per):
            return (t.data_ptr() + 8 * per * step).contiguous()

        info = torch.zeros(count, dtype=torch.int32, device=dev)
        hip.dposv_backward_vbatched(ns, ptrs(l, n * n), ns, ptrs(x, n), ptrs(xbar, n), ptrs(abar, n * n), ns,
                                    ptrs(bbar, n), info, n)
        # abar is symmetric: its column-major image is the same row-major tensor
        return abar, bbar


def spd_solve(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Batched x = S^-1 b, S = (A + A^T)/2, differentiable in a and b."""
    return SpdSolve.apply(a, b)


class Cholesky(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor):
        if a.dtype != torch.float64 or not a.is_cuda:
            raise TypeError("cholesky works on float64 CUDA (HIP) tensors: there is no CPU path")
        if a.dim() != 3 or a.shape[1] != a.shape[2] or not (1 <= a.shape[1] <= hip.SMALL_MAX_N):
            raise ValueError(f"a must have shape (count, n, n) with n in [1, {hip.SMALL_MAX_N}]")
        count, n = a.shape[0], a.shape[1]
        # S = (A + A^T)/2 stored row-major is its own column-major image; the
        # kernel overwrites it with L in column-major order, so the row-major
        # tensor then holds L^T.
        buf = (0.5 * (a + a.transpose(1, 2))).contiguous()
        dev = a.device
        ns = torch.full((count,), n, dtype=torch.int32, device=dev)
        step = torch.arange(count, dtype=torch.int64, device=dev)
        ptrs = (buf.data_ptr() + 8 * n * n * step).contiguous()
        info = torch.zeros(count, dtype=torch.int32, device=dev)
        h = hip.default_handle()
        mode = h.math_mode()
        h.set_math_mode(hip.MATH_EXACT)
        try:
            hip.dpotrf_vbatched(ns, ptrs, ns, info, n, handle=h)
        finally:
            h.set_math_mode(mode)
        bad = torch.nonzero(info).flatten()
        if bad.numel():
            k = int(bad[0].item())
            raise hip.MatrixFactorizationException(
                f"Cholesky decomposition failed for matrix {k} at row {int(info[k].item()) - 1}. "
                "The matrix must be positive definite.")
        ctx.save_for_backward(buf)
        return buf.transpose(1, 2)

    @staticmethod
    def backward(ctx, lbar: torch.Tensor):
        (buf,) = ctx.saved_tensors
        count, n = buf.shape[0], buf.shape[1]
        dev = buf.device
        # column-major images: buf already is L's; lbar's is its transpose, stored row-major
        lbar_cm = lbar.transpose(1, 2).contiguous()
        abar = torch.empty_like(buf)
        ns = torch.full((count,), n, dtype=torch.int32, device=dev)
        step = torch.arange(count, dtype=torch.int64, device=dev)

        def ptrs(t):
            return (t.data_ptr() + 8 * n * n * step).contiguous()

        info = torch.zeros(count, dtype=torch.int32, device=dev)
        hip.dpotrf_backward_vbatched(ns, ptrs(buf), ns, ptrs(lbar_cm), ns, ptrs(abar), ns, info, n)
        return abar  # symmetric: its column-major image is the same tensor


def cholesky(a: torch.Tensor) -> torch.Tensor:
    """Batched lower Cholesky factor of S = (A + A^T)/2, differentiable in a."""
    return Cholesky.apply(a)
