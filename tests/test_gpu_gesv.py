"""GPU tests of the blocked LU solvers for n <= 512: the plain FP64 path
(mfh_dgesv_vbatched) and the mixed-precision path (FP32 factorization with
BF16-split MFMA trailing updates at n >= 256, FP64 iterative refinement;
mfh_dgesv_mixed_vbatched).

Neither path has a bit-identical MatFlat counterpart: they compose
Factorization.Lu (Factorization.Lu.cs:122-161) with the LU solve of
Factorization.LuInverse.cs:169-192 and are judged by the normwise backward
error eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), recomputed
here on the CPU in FP64 from the returned x, and by the forward error against
LAPACK (numpy.linalg.solve).  Test matrices are A = U diag(s) V^T with
Haar-random U, V and log-spaced singular values, as in the mixed-precision
workload.
"""
import numpy as np
import pytest
import torch

from gpu_util import Packed, assert_bits_equal, f64, i32
from helpers import backward_error, view

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def hipmod():
    from matflat_amd import hip
    return hip


def conditioned(seed: int, n: int, cond: float) -> np.ndarray:
    rng = np.random.default_rng(seed)

    def haar():
        q, r = np.linalg.qr(rng.standard_normal((n, n)))
        return q * np.sign(np.diag(r))

    s = cond ** (-np.arange(n) / max(n - 1, 1))
    return (haar() * s) @ haar().T


def col_major(a: np.ndarray, lda: int) -> np.ndarray:
    n = a.shape[0]
    buf = np.full(lda * n, np.nan)
    for j in range(n):
        buf[j * lda:j * lda + n] = a[:, j]
    return buf


CASES = [(1, 1, 1e1), (2, 3, 1e2), (5, 5, 1e3), (16, 16, 1e4), (31, 40, 1e2), (33, 33, 1e6), (64, 64, 1e5),
         (100, 104, 1e6), (128, 128, 1e6), (130, 131, 1e3), (255, 256, 1e4), (256, 256, 1e6), (257, 260, 1e2),
         (300, 300, 1e5), (384, 384, 1e6), (512, 512, 1e6)]


def build(cases, seed0):
    mats, bufs, bs = [], [], []
    for i, (n, lda, cond) in enumerate(cases):
        a = conditioned(seed0 + i, n, cond)
        mats.append(a)
        bufs.append(col_major(a, lda))
        bs.append(np.random.default_rng(seed0 + 500 + i).standard_normal(n))
    return mats, bufs, bs


@pytest.mark.parametrize("mixed", [False, True])
def test_gesv_backward_and_forward_error(mixed):
    hip = hipmod()
    mats, bufs, bs = build(CASES, 11 if mixed else 17)
    ns = [c[0] for c in CASES]
    ldas = [c[1] for c in CASES]
    A, B, X = Packed(bufs), Packed(bs), Packed([np.zeros(n) for n in ns])
    count = len(ns)
    info = i32(np.full(count, 99))
    berr = f64(np.full(count, -1.0))
    iters = i32(np.full(count, -1))
    if mixed:
        hip.dgesv_mixed_vbatched(i32(ns), A.ptrs, i32(ldas), B.ptrs, X.ptrs, 10, 1e-15, iters, berr, info, 512)
    else:
        hip.dgesv_vbatched(i32(ns), A.ptrs, i32(ldas), B.ptrs, X.ptrs, berr, info, 512)
    torch.cuda.synchronize()
    got_info, got_berr, got_iters = info.cpu().numpy(), berr.cpu().numpy(), iters.cpu().numpy()
    xs, a_after, b_after = X.all(), A.all(), B.all()
    for i, (n, lda, cond) in enumerate(CASES):
        assert got_info[i] == 0, (i, n, got_info[i])
        eta = backward_error(mats[i], xs[i], bs[i])
        bound = 1e-15 * n if mixed else max(n, 8) * EPS
        assert eta <= bound, (i, n, cond, eta)
        # the device's own estimate agrees with the CPU recomputation
        assert abs(got_berr[i] - eta) <= 0.5 * eta + 4 * EPS, (i, got_berr[i], eta)
        want = np.linalg.solve(mats[i], bs[i])
        fwd = np.abs(xs[i] - want).max() / np.abs(want).max()
        assert fwd <= cond * n * 20 * EPS, (i, n, cond, fwd)
        if mixed:
            assert 0 <= got_iters[i] <= 10
        # A and b are not modified
        assert_bits_equal(a_after[i], bufs[i])
        assert_bits_equal(b_after[i], bs[i])


def test_gesv_mixed_needs_refinement_at_high_condition():
    """An FP32 factorization alone cannot reach 1e-15*n at cond 1e6: the
    refinement steps are taken (iters >= 1) and the target is met."""
    hip = hipmod()
    cases = [(256, 256, 1e6), (512, 512, 1e6)]
    mats, bufs, bs = build(cases, 91)
    ns = [c[0] for c in cases]
    A, B, X = Packed(bufs), Packed(bs), Packed([np.zeros(n) for n in ns])
    info, berr, iters = i32([99, 99]), f64([-1.0, -1.0]), i32([-1, -1])
    hip.dgesv_mixed_vbatched(i32(ns), A.ptrs, i32(ns), B.ptrs, X.ptrs, 10, 1e-15, iters, berr, info, 512)
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    assert np.all(iters.cpu().numpy() >= 1)
    xs = X.all()
    for i in range(2):
        assert backward_error(mats[i], xs[i], bs[i]) <= 1e-15 * ns[i]


@pytest.mark.parametrize("mixed", [False, True])
def test_gesv_singular_and_invalid(mixed):
    """A zero column gives info = (first zero pivot) + 1; an order above
    max_n gives info = -3 and leaves x untouched; neighbours are solved."""
    hip = hipmod()
    n = 40
    good = conditioned(5, n, 1e2)
    sing = good.copy()
    sing[:, 7] = 0.0
    big = 24
    bufs = [col_major(good, n), col_major(sing, n), col_major(conditioned(6, big, 10.0), big)]
    bs = [np.random.default_rng(k).standard_normal(m) for k, m in enumerate([n, n, big])]
    A, B = Packed(bufs), Packed(bs)
    X = Packed([np.full(n, 7.0), np.full(n, 7.0), np.full(big, 7.0)])
    info, berr, iters = i32([99] * 3), f64([-1.0] * 3), i32([-1] * 3)
    max_n = 40
    if mixed:
        # the third matrix claims an order above max_n
        hip.dgesv_mixed_vbatched(i32([n, n, 41]), A.ptrs, i32([n, n, 41]), B.ptrs, X.ptrs, 10, 1e-15, iters, berr,
                                 info, max_n)
    else:
        hip.dgesv_vbatched(i32([n, n, 41]), A.ptrs, i32([n, n, 41]), B.ptrs, X.ptrs, berr, info, max_n)
    torch.cuda.synchronize()
    got = info.cpu().numpy()
    assert got[0] == 0
    assert got[1] == 8, got  # column 7 is zero: the pivot at step 7 is zero
    assert got[2] == -3
    xs = X.all()
    assert backward_error(good, xs[0], bs[0]) <= 1e-15 * n
    assert np.all(xs[2] == 7.0)


def test_gesv_argument_errors():
    hip = hipmod()
    one = i32([4])
    ptr = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(hip.ArgumentError) as e:
        hip.dgesv_vbatched(one, ptr, one, ptr, ptr, f64([0.0]), one, 513)
    assert e.value.position == 10
    with pytest.raises(hip.ArgumentError) as e:
        hip.dgesv_mixed_vbatched(one, ptr, one, ptr, ptr, 10, 0.0, one, f64([0.0]), one, 8)
    assert e.value.position == 9


@pytest.mark.parametrize("mixed", [False, True])
def test_gesv_nan_and_fp32_range(mixed):
    """A NaN in A gives a NaN backward error, reported as a failure (info =
    n+1 on the FP64 path), never as convergence.  An entry beyond the FP32
    range cannot be factored in FP32: the mixed path reports info = n+2,
    iters = -2 and leaves x untouched (LAPACK dsgesv's dlag2s check); the
    FP64 path solves the same matrix to backward error <= n eps.  The
    neighbours in the batch are solved."""
    hip = hipmod()
    n = 48
    good = conditioned(21, n, 1e3)
    nanm = good.copy()
    nanm[5, 9] = np.nan
    big = good * 1e41  # entries of this one are beyond the FP32 range (3.4e38)
    mats = [good, nanm, big, good]
    bufs = [col_major(m, n) for m in mats]
    bs = [np.random.default_rng(70 + k).standard_normal(n) for k in range(4)]
    bs[2] = bs[2] * 1e41
    A, B = Packed(bufs), Packed(bs)
    X = Packed([np.full(n, 7.0) for _ in range(4)])
    info, berr, iters = i32([99] * 4), f64([-1.0] * 4), i32([-1] * 4)
    if mixed:
        hip.dgesv_mixed_vbatched(i32([n] * 4), A.ptrs, i32([n] * 4), B.ptrs, X.ptrs, 10, 1e-15, iters, berr, info, n)
    else:
        hip.dgesv_vbatched(i32([n] * 4), A.ptrs, i32([n] * 4), B.ptrs, X.ptrs, berr, info, n)
    torch.cuda.synchronize()
    got, be, it, xs = info.cpu().numpy(), berr.cpu().numpy(), iters.cpu().numpy(), X.all()
    for k in (0, 3):
        assert got[k] == 0, got
        assert backward_error(good, xs[k], bs[k]) <= (1e-15 * n if mixed else n * EPS)
    if mixed:
        # the NaN entry is caught by the same range test (it is not finite)
        assert got[1] == n + 2 and it[1] == -2, (got, it)
        assert got[2] == n + 2 and it[2] == -2, (got, it)
        assert np.all(xs[1] == 7.0) and np.all(xs[2] == 7.0)
    else:
        assert got[1] == n + 1, got
        assert np.isnan(be[1])
        assert got[2] == 0, got
        assert backward_error(big, xs[2], bs[2]) <= n * EPS


def test_getrf_blocked_factors_and_pivots():
    """mfh_dgetrf_blocked_vbatched factors in place with MatFlat's piv
    convention ((PA)[i, :] = A[piv[i], :], Factorization.Lu.cs:296-350):
    piv is a permutation, |L| <= 1, ||PA - LU||_1 / (||A||_1 n eps) <= 10
    (LAPACK dget01), padding rows untouched; a zero column gives info =
    (first zero pivot) + 1 and the factorization still completes."""
    hip = hipmod()
    cases = [(1, 1, 1e1), (7, 9, 1e2), (64, 64, 1e4), (129, 131, 1e3), (256, 256, 1e5), (383, 384, 1e6),
             (512, 512, 1e6)]
    mats, bufs, _ = build(cases, 301)
    sing = conditioned(9, 40, 1e2)
    sing[:, 11] = 0.0
    mats.append(sing)
    bufs.append(col_major(sing, 40))
    ns = [c[0] for c in cases] + [40]
    ldas = [c[1] for c in cases] + [40]
    A = Packed(bufs)
    P = Packed([np.full(n, -1, dtype=np.int32) for n in ns], dtype=torch.int32)
    info = i32(np.full(len(ns), 99))
    hip.dgetrf_blocked_vbatched(i32(ns), A.ptrs, i32(ldas), P.ptrs, info, 512)
    torch.cuda.synchronize()
    got, ga, gp = info.cpu().numpy(), A.all(), P.all()
    for i, (n, lda) in enumerate(zip(ns, ldas)):
        piv = gp[i]
        assert sorted(piv.tolist()) == list(range(n)), i
        lu = view(ga[i], n, n, lda)
        low = np.tril(lu, -1) + np.eye(n)
        up = np.triu(lu)
        # pivots are the largest |a| to 2^-20 relative
        assert np.all(np.abs(np.tril(lu, -1)) <= 1.0 + 1e-5), i
        pa = mats[i][piv, :]
        ratio = np.abs(pa - low @ up).sum(axis=0).max() / (np.abs(mats[i]).sum(axis=0).max() * n * EPS)
        assert ratio <= 10.0, (i, n, ratio)
        for j in range(n):
            assert np.all(np.isnan(ga[i][j * lda + n:(j + 1) * lda])), i
        if i == len(ns) - 1:
            assert got[i] == 12, got
        else:
            assert got[i] == 0, (i, got[i])
