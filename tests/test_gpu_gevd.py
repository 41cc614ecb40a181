"""Factorization.Gevd (double) on the GPU (mfh_dgevd_vbatched): eigenvectors
in a, the Cholesky factor in b, eigenvalues in w and info must equal the
oracle's (oracle/matflat_oracle.c: mfo_dgevd, a restatement of
Factorization.Gevd.cs:330-376 with SvdCore, Factorization.Svd.cs:786-1256)
bit for bit, through the C ABI, for both LDS layouts (U in LDS up to order
100, in a above), with NaN lower triangles, failures and invalid arguments."""
import os

import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, i32
from helpers import gevd_pair_bufs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hipmod():
    from matflat_amd import hip
    return hip


def run(cases, a_bufs, b_bufs, max_n, math="exact"):
    """cases: (n, lda, ldb); returns (a, b, w, info) lists from the GPU."""
    hip = hipmod()
    h = hip.Handle()
    h.set_math_mode(hip.MATH_EXACT if math == "exact" else hip.MATH_FAST)
    h.bind_stream()
    A, B = Packed(a_bufs), Packed(b_bufs)
    W = Packed([np.full(n, -3.0) for n, _, _ in cases])
    info = i32(np.full(len(cases), 99))
    ns = i32([c[0] for c in cases])
    hip.dgevd_vbatched(ns, A.ptrs, i32([c[1] for c in cases]), B.ptrs, i32([c[2] for c in cases]), W.ptrs, info,
                       max_n, handle=h)
    torch.cuda.synchronize()
    out = A.all(), B.all(), W.all(), info.cpu().numpy()
    h.close()
    return out


def oracle_run(n, lda, ldb, a_buf, b_buf):
    a, b = a_buf.copy(), b_buf.copy()
    w = np.full(n, -3.0)
    info = oracle.dgevd(n, a, lda, b, ldb, w)
    return a, b, w, info


def check(cases, a_bufs, b_bufs, max_n, math="exact"):
    ga, gb, gw, gi = run(cases, a_bufs, b_bufs, max_n, math)
    for i, (n, lda, ldb) in enumerate(cases):
        wa, wb, ww, wi = oracle_run(n, lda, ldb, a_bufs[i], b_bufs[i])
        assert gi[i] == wi, (i, n, gi[i], wi)
        assert_bits_equal(gb[i], wb, f"gevd b {i} n={n}")
        assert_bits_equal(ga[i], wa, f"gevd a {i} n={n}")
        assert_bits_equal(gw[i], ww, f"gevd w {i} n={n}")
    return gi


@pytest.mark.parametrize("small", [True, False])
def test_golden_gevd(small):
    """The committed fixture (GevdTests.cs:94-105 shapes and larger ones):
    small runs the orders up to 100 (max_n 100); otherwise all orders with
    max_n 128."""
    hip = hipmod()
    g = np.load(os.path.join(GOLDEN, "gevd.npz"), allow_pickle=False)
    meta = g["gevd_meta"]
    sel = [i for i, row in enumerate(meta) if row[0] <= 100 or not small]
    cases = [tuple(int(v) for v in meta[i][:3]) for i in sel]
    a_in = [g["gevd_a_in"][g["gevd_a_off"][i]:g["gevd_a_off"][i] + meta[i][1] * meta[i][0]] for i in sel]
    b_in = [g["gevd_b_in"][g["gevd_b_off"][i]:g["gevd_b_off"][i] + meta[i][2] * meta[i][0]] for i in sel]
    max_n = max(c[0] for c in cases)
    ga, gb, gw, gi = run(cases, a_in, b_in, max_n)
    for k, i in enumerate(sel):
        n, lda, ldb, info = (int(v) for v in meta[i])
        assert gi[k] == info, (k, gi[k], info)
        ao, bo, wo = g["gevd_a_off"][i], g["gevd_b_off"][i], g["gevd_w_off"][i]
        assert_bits_equal(ga[k], g["gevd_a_out"][ao:ao + lda * n], f"gevd a {k}")
        assert_bits_equal(gb[k], g["gevd_b_out"][bo:bo + ldb * n], f"gevd b {k}")
        if info == 0:
            assert_bits_equal(gw[k], g["gevd_w"][wo:wo + n], f"gevd w {k}")
        else:
            assert np.all(gw[k] == -3.0)
    del hip


@pytest.mark.parametrize("max_n", [64, 100, 128])
def test_gevd_variable_sizes(max_n):
    rng = np.random.default_rng(max_n)
    count = 96 if max_n <= 64 else 40
    cases = [(int(n), int(n + rng.integers(0, 4)), int(n + rng.integers(0, 4)))
             for n in rng.integers(1, max_n + 1, size=count)]
    cases[0] = (max_n, max_n, max_n + 1)
    a_bufs, b_bufs = [], []
    for i, (n, lda, ldb) in enumerate(cases):
        a, b, _, _ = gevd_pair_bufs(1000 * max_n + i, n, lda, ldb, shift=0.25)
        a_bufs.append(a)
        b_bufs.append(b)
    gi = check(cases, a_bufs, b_bufs, max_n)
    assert np.all(gi == 0)


def test_gevd_fast_mode_handle_is_still_exact():
    cases = [(n, n + 1, n) for n in (3, 17, 40)]
    bufs = [gevd_pair_bufs(77 + i, *c, shift=0.25) for i, c in enumerate(cases)]
    check(cases, [b[0] for b in bufs], [b[1] for b in bufs], 40, math="fast")


def test_gevd_failures():
    """B not positive definite at rows 0, 3 and n-1 (info = j+1, b holds the
    Cholesky's partial writes, a and w untouched) beside good matrices; a NaN
    in A's upper triangle (the iteration does not converge: info = n+1, a holds
    the vectors reached, w untouched)."""
    cases, a_bufs, b_bufs = [], [], []
    for i, (n, row) in enumerate([(6, 3), (5, 0), (9, 8), (7, None), (8, -1)]):
        a, b, _, _ = gevd_pair_bufs(300 + i, n, n + 1, n + 2, shift=0.25)
        if row is not None and row >= 0:
            b[row + row * (n + 2)] = -1e3
        if row == -1:
            a[2 + 5 * (n + 1)] = np.nan
        cases.append((n, n + 1, n + 2))
        a_bufs.append(a)
        b_bufs.append(b)
    gi = check(cases, a_bufs, b_bufs, 9)
    assert list(gi) == [4, 1, 9, 0, 9]


def test_gevd_invalid_arguments():
    """Per-matrix argument errors: -(position) in info, nothing written."""
    hip = hipmod()
    good = gevd_pair_bufs(9, 4, 4, 4, shift=0.25)
    cases = [(4, 4, 4), (0, 4, 4), (4, 3, 4), (4, 4, 3), (5, 5, 5), (4, 4, 4), (4, 4, 4), (4, 4, 4)]
    A = Packed([good[0]] * len(cases))
    B = Packed([good[1]] * len(cases))
    W = Packed([np.full(5, -3.0)] * len(cases))
    aptr, bptr, wptr = A.ptrs.clone(), B.ptrs.clone(), W.ptrs.clone()
    aptr[5] = 0
    bptr[6] = 0
    wptr[7] = 0
    info = i32(np.full(len(cases), 99))
    h = hip.Handle()
    h.bind_stream()
    hip.dgevd_vbatched(i32([c[0] for c in cases]), aptr, i32([c[1] for c in cases]), bptr, i32([c[2] for c in cases]),
                       wptr, info, 4, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert list(info.cpu().numpy()) == [0, -3, -5, -7, -3, -4, -6, -8]
    ga, gb, gw = A.all(), B.all(), W.all()
    for i in range(1, len(cases)):
        assert_bits_equal(ga[i], good[0], f"a {i}")
        assert_bits_equal(gb[i], good[1], f"b {i}")
        assert np.all(gw[i] == -3.0)
    with pytest.raises(hip.ArgumentError):
        hip.dgevd_vbatched(i32([4]), A.ptrs[:1], i32([4]), B.ptrs[:1], i32([4]), W.ptrs[:1], info[:1], 129)


def test_dropin_gevd():
    """Factorization.Gevd with the reference's signature, on GevdTests.cs's
    shapes: bit-identical to the oracle, and its exceptions."""
    from matflat_amd import matflat
    for i, (n, lda, ldb) in enumerate([(1, 1, 1), (2, 4, 3), (5, 7, 8), (10, 17, 11)]):
        a_buf, b_buf, _, _ = gevd_pair_bufs(500 + i, n, lda, ldb)
        a = torch.from_numpy(a_buf.copy()).cuda()
        b = torch.from_numpy(b_buf.copy()).cuda()
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        matflat.Factorization.Gevd(n, a, lda, b, ldb, w)
        wa, wb, ww, wi = oracle_run(n, lda, ldb, a_buf, b_buf)
        assert wi == 0
        assert_bits_equal(a.cpu().numpy(), wa, "a")
        assert_bits_equal(b.cpu().numpy(), wb, "b")
        assert_bits_equal(w.cpu().numpy(), ww[:n], "w")
    a_buf, b_buf, _, _ = gevd_pair_bufs(600, 4, 4, 4)
    b_buf[0] = -1.0
    a = torch.from_numpy(a_buf.copy()).cuda()
    b = torch.from_numpy(b_buf.copy()).cuda()
    w = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(matflat.MatFlatException) as e:
        matflat.Factorization.Gevd(4, a, 4, b, 4, w)
    assert not isinstance(e.value, matflat.MatrixFactorizationException)
    assert "right-hand side matrix is not positive definite" in str(e.value)
    with pytest.raises(matflat.ArgumentException):
        matflat.Factorization.Gevd(4, a, 3, b, 4, w)
    with pytest.raises(matflat.ArgumentException):
        matflat.Factorization.Gevd(4, a, 4, None, 4, w)
