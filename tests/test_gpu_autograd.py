"""Gradient checks of the SPD-solve adjoint through matflat_amd.autograd.

The workload is 50,000 SPD matrices of order 32: A = B B^T + 32 I, with b and
w ~ N(0,1) and loss = w . x. It is checked here on a batch small enough for
finite differences. The adjoint kernel (mfh_dposv_backward_vbatched) has no
MatFlat counterpart.

The checks are:
- relative error <= 1e-8 against central finite differences with h = 1e-6,
  for symmetric perturbations of A and for b;
- agreement with the CPU reference gradient of the oracle;
- torch.autograd.gradcheck.
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def spd_batch(count, n, seed):
    g = torch.Generator().manual_seed(seed)
    bm = torch.randn((count, n, n), dtype=torch.float64, generator=g)
    a = bm @ bm.transpose(1, 2).contiguous() + n * torch.eye(n, dtype=torch.float64)
    b = torch.randn((count, n), dtype=torch.float64, generator=g)
    w = torch.randn((count, n), dtype=torch.float64, generator=g)
    return a.cuda(), b.cuda(), w.cuda()


def loss_of(a, b, w):
    from matflat_amd.autograd import spd_solve
    return (w * spd_solve(a, b)).sum()


def test_gradient_vs_central_differences_n32():
    """Normwise relative error of the analytic gradient against central
    differences (h = 1e-6) over all of b and 64 symmetric pairs of A."""
    a, b, w = spd_batch(2, 32, 1)
    a.requires_grad_(True)
    b.requires_grad_(True)
    ga, gb = torch.autograd.grad(loss_of(a, b, w), (a, b))
    h = 1e-6
    rng = np.random.default_rng(2)
    fd, an = [], []
    with torch.no_grad():
        for i in range(32):
            d = torch.zeros_like(b)
            d[0, i] = h
            fd.append(float((loss_of(a, b + d, w) - loss_of(a, b - d, w)) / (2 * h)))
            an.append(float(gb[0, i]))
        for _ in range(64):
            i, j = sorted(int(v) for v in rng.integers(0, 32, size=2))
            e = torch.zeros_like(a)
            e[0, i, j] = h
            e[0, j, i] = h
            fd.append(float((loss_of(a + e, b, w) - loss_of(a - e, b, w)) / (2 * h)))
            an.append(float(ga[0, i, j] + ga[0, j, i]) if i != j else float(ga[0, i, i]))
    fd, an = np.array(fd), np.array(an)
    rel = np.linalg.norm(fd - an) / np.linalg.norm(an)
    assert rel <= 1e-8, rel


def test_gradient_matches_cpu_reference():
    count, n = 6, 32
    a, b, w = spd_batch(count, n, 3)
    a.requires_grad_(True)
    b.requires_grad_(True)
    ga, gb = torch.autograd.grad(loss_of(a, b, w), (a, b))
    ga, gb = ga.cpu().numpy(), gb.cpu().numpy()
    ah, bh, wh = a.detach().cpu().numpy(), b.detach().cpu().numpy(), w.cpu().numpy()
    for k in range(count):
        sym = 0.5 * (ah[k] + ah[k].T)  # the forward factors the symmetric part, computed the same way
        buf = sym.T.reshape(-1).copy()  # column-major
        x = bh[k].copy()
        assert oracle.dposv(n, buf, n, x) == 0
        abar, bbar = oracle.dposv_backward(n, buf, n, x, wh[k])
        # the oracle is the same restatement of MatFlat's order: bit-identical
        assert np.array_equal(gb[k], bbar)
        assert np.array_equal(ga[k].T.reshape(-1), abar)


def test_torch_gradcheck_small():
    a, b, _ = spd_batch(2, 6, 5)
    a.requires_grad_(True)
    b.requires_grad_(True)
    from matflat_amd.autograd import spd_solve
    assert torch.autograd.gradcheck(spd_solve, (a, b), eps=1e-6, atol=1e-8, rtol=1e-7)


def test_not_positive_definite_raises():
    from matflat_amd import hip
    a, b, w = spd_batch(2, 8, 7)
    a[1, 3, 3] = -1e3
    with pytest.raises(hip.MatrixFactorizationException):
        loss_of(a, b, w)
