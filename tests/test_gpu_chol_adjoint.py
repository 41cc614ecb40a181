"""The adjoint of the batched Cholesky factorization on the GPU
(mfh_dpotrf_backward_vbatched; no MatFlat counterpart): bit-identical to the
oracle (oracle/matflat_oracle.c: mfo_dpotrf_backward, pinned by finite
differences in tests/test_oracle_chol_adjoint.py) over variable orders and
leading dimensions in both LDS layouts, per-matrix argument errors, and
through matflat_amd.autograd.cholesky: central differences (<= 1e-8
relative, h = 1e-6) and torch.autograd.gradcheck."""
import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, i32

pytestmark = pytest.mark.gpu


def hipmod():
    from matflat_amd import hip
    return hip


def case(seed, n, ldl, ldlb, ldab):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    a = b @ b.T + n * np.eye(n)
    l = np.linalg.cholesky(a)
    lb = np.tril(rng.standard_normal((n, n)))
    lbuf = np.full(ldl * n, np.nan)
    lbbuf = np.full(ldlb * n, np.nan)
    for j in range(n):
        lbuf[j * ldl:j * ldl + n] = l[:, j]
        lbbuf[j * ldlb:j * ldlb + n] = lb[:, j]
        lbbuf[j * ldlb:j * ldlb + j] = np.nan  # strict upper of lbar: never read
    lbuf_clean = lbuf.copy()
    for j in range(n):
        lbuf_clean[j * ldl:j * ldl + j] = 0.0  # as the Cholesky leaves it
    return lbuf_clean, lbbuf


@pytest.mark.parametrize("max_n", [32, 100, 128])
def test_matches_oracle_variable_sizes(max_n):
    hip = hipmod()
    rng = np.random.default_rng(max_n)
    count = 120 if max_n <= 64 else 40
    dims = []
    for k in range(count):
        n = max_n if k == 0 else int(rng.integers(1, max_n + 1))
        dims.append((n, n + int(rng.integers(0, 3)), n + int(rng.integers(0, 3)), n + int(rng.integers(0, 3))))
    ls, lbs = zip(*[case(1000 * max_n + k, *d) for k, d in enumerate(dims)])
    L, LB = Packed(list(ls)), Packed(list(lbs))
    AB = Packed([np.full(d[3] * d[0], -7.0) for d in dims])
    info = i32(np.full(count, 99))
    h = hip.Handle()
    h.bind_stream()
    hip.dpotrf_backward_vbatched(i32([d[0] for d in dims]), L.ptrs, i32([d[1] for d in dims]), LB.ptrs,
                                 i32([d[2] for d in dims]), AB.ptrs, i32([d[3] for d in dims]), info, max_n, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.all(info.cpu().numpy() == 0)
    got = AB.all()
    for k, (n, ldl, ldlb, ldab) in enumerate(dims):
        want = np.full(ldab * n, -7.0)
        oracle.lib().mfo_dpotrf_backward(n, ls[k].ctypes.data, ldl, lbs[k].ctypes.data, ldlb, want.ctypes.data, ldab)
        assert_bits_equal(got[k], want, f"abar {k} n={n}")


def test_invalid_arguments():
    hip = hipmod()
    l, lb = case(5, 4, 4, 4, 4)
    dims = [(4, 4, 4, 4), (0, 4, 4, 4), (4, 3, 4, 4), (4, 4, 3, 4), (4, 4, 4, 3), (5, 5, 5, 5), (4, 4, 4, 4),
            (4, 4, 4, 4), (4, 4, 4, 4)]
    L, LB = Packed([l] * len(dims)), Packed([lb] * len(dims))
    AB = Packed([np.full(16, -7.0)] * len(dims))
    pl, plb, pab = L.ptrs.clone(), LB.ptrs.clone(), AB.ptrs.clone()
    pl[6] = 0
    plb[7] = 0
    pab[8] = 0
    info = i32(np.full(len(dims), 99))
    cols = list(zip(*dims))
    hip.dpotrf_backward_vbatched(i32(cols[0]), pl, i32(cols[1]), plb, i32(cols[2]), pab, i32(cols[3]), info, 4)
    torch.cuda.synchronize()
    assert list(info.cpu().numpy()) == [0, -3, -5, -7, -9, -3, -4, -6, -8]
    got = AB.all()
    for k in range(1, len(dims)):
        assert np.all(got[k] == -7.0)


def chol_loss(a, w):
    from matflat_amd.autograd import cholesky
    return (w * cholesky(a)).sum()


def test_autograd_cholesky_forward_is_matflat():
    """The forward returns L bit-identical to the exact Cholesky."""
    from matflat_amd.autograd import cholesky
    g = torch.Generator().manual_seed(3)
    bm = torch.randn((4, 9, 9), dtype=torch.float64, generator=g)
    a = bm @ bm.transpose(1, 2) + 9 * torch.eye(9, dtype=torch.float64)
    l = cholesky(a.cuda()).cpu().numpy()
    for k in range(4):
        s = (0.5 * (a[k] + a[k].T)).numpy()
        buf = s.T.reshape(-1).copy()
        assert oracle.dpotrf(9, buf, 9) == 0
        assert_bits_equal(l[k], buf.reshape(9, 9).T, f"L {k}")


def test_autograd_cholesky_central_differences():
    n, count = 16, 3
    g = torch.Generator().manual_seed(11)
    bm = torch.randn((count, n, n), dtype=torch.float64, generator=g)
    a = (bm @ bm.transpose(1, 2) + n * torch.eye(n, dtype=torch.float64)).cuda()
    w = torch.tril(torch.randn((count, n, n), dtype=torch.float64, generator=g)).cuda()
    a.requires_grad_(True)
    chol_loss(a, w).backward()
    grad = a.grad.detach()
    h = 1e-6
    rng = np.random.default_rng(2)
    with torch.no_grad():
        for _ in range(24):
            k, i, j = int(rng.integers(count)), int(rng.integers(n)), int(rng.integers(n))
            e = torch.zeros_like(a)
            e[k, i, j] += 1.0
            if i != j:
                e[k, j, i] += 1.0
            fd = (chol_loss(a + h * e, w) - chol_loss(a - h * e, w)).item() / (2 * h)
            # the loss sees S = (A + A^T)/2: a symmetric pair moves S[i, j] by one
            pred = float((grad * e).sum().item()) if i == j else float(2 * grad[k, i, j].item())
            assert abs(fd - pred) <= 1e-8 * max(1.0, abs(fd)) + 1e-9, (k, i, j, fd, pred)


def test_autograd_cholesky_gradcheck():
    from matflat_amd.autograd import cholesky
    g = torch.Generator().manual_seed(5)
    bm = torch.randn((2, 5, 5), dtype=torch.float64, generator=g)
    a = (bm @ bm.transpose(1, 2) + 5 * torch.eye(5, dtype=torch.float64)).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: torch.tril(cholesky(x)), (a,), eps=1e-6, atol=1e-8, rtol=1e-7)
