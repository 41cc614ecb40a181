"""The HIP exact path against the committed golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): LU factors, pivots and signs, Cholesky factors, SPD-solve x, the
eight triangular-solve forms, GEMM (real and complex, every transpose /
conjugate combination), GEMV and the LU inverse must match bit for bit.  Shapes follow the
reference's tests (LuTests.cs:103-122, CholeskyTests.cs,
BlasTests.SolveTriangular*.cs, BlasTests.MulMatMat*.cs)."""
import os

import numpy as np
import pytest
import torch

from gpu_util import Packed, assert_bits_equal, f64, i32

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def hipmod():
    from matflat_amd import hip
    return hip


def exact_handle():
    hip = hipmod()
    h = hip.Handle()
    h.set_math_mode(hip.MATH_EXACT)
    return h


def test_golden_lu():
    hip = hipmod()
    g = load("lu.npz")
    meta = g["lu_meta"]
    ms, ns, ldas = meta[:, 0], meta[:, 1], meta[:, 2]
    bufs = [g["lu_in"][o:o + l * n].copy() for o, l, n in zip(g["lu_off"], ldas, ns)]
    A = Packed(bufs)
    P = Packed([np.full(m, -7, dtype=np.int32) for m in ms], dtype=torch.int32)
    sign = i32(np.zeros(len(ms)))
    info = i32(np.full(len(ms), 99))
    h = exact_handle()
    hip.dgetrf_vbatched(i32(ms), i32(ns), A.ptrs, i32(ldas), P.ptrs, sign, info, int(ms.max()), int(ns.max()),
                        handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.all(info.cpu().numpy() == 0)
    ga, gp, gs = A.all(), P.all(), sign.cpu().numpy()
    for i, (m, n, lda, sg) in enumerate(meta):
        o, po = g["lu_off"][i], g["lu_piv_off"][i]
        assert_bits_equal(ga[i], g["lu_out"][o:o + lda * n], f"lu {i} {(m, n, lda)}")
        assert np.array_equal(gp[i], g["lu_piv"][po:po + m]), i
        assert gs[i] == sg, i


def test_golden_cholesky_and_posv():
    hip = hipmod()
    g = load("cholesky_posv.npz")
    meta = g["chol_meta"]
    ns, ldas = meta[:, 0], meta[:, 1]
    ins = [g["chol_in"][o:o + l * n].copy() for o, l, n in zip(g["chol_off"], ldas, ns)]
    h = exact_handle()
    A = Packed(ins)
    info = i32(np.full(len(ns), 99))
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, int(ns.max()), handle=h)
    A2 = Packed(ins)
    X = Packed([g["posv_b"][o:o + n].copy() for o, n in zip(g["posv_off"], ns)])
    info2 = i32(np.full(len(ns), 99))
    hip.dposv_vbatched(i32(ns), A2.ptrs, i32(ldas), X.ptrs, info2, int(ns.max()), handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.array_equal(info.cpu().numpy(), meta[:, 2])
    assert np.all(info2.cpu().numpy() == 0)
    ga, gx = A.all(), X.all()
    for i, (n, lda, _) in enumerate(meta):
        o, bo = g["chol_off"][i], g["posv_off"][i]
        assert_bits_equal(ga[i], g["chol_out"][o:o + lda * n], f"potrf n={n}")
        assert_bits_equal(gx[i], g["posv_x"][bo:bo + n], f"posv n={n}")


def test_golden_trsv():
    hip = hipmod()
    g = load("trsv.npz")
    meta = g["trsv_meta"]
    h = exact_handle()
    for form in sorted({tuple(r[:3]) for r in meta}):
        sel = [i for i, r in enumerate(meta) if tuple(r[:3]) == form]
        ns = [int(meta[i, 3]) for i in sel]
        ldas = [int(meta[i, 4]) for i in sel]
        incs = [int(meta[i, 5]) for i in sel]
        A = Packed([g["trsv_a"][g["trsv_off"][i]:g["trsv_off"][i] + meta[i, 4] * meta[i, 3]].copy() for i in sel])
        X = Packed([g["trsv_x_in"][g["trsv_x_off"][i]:g["trsv_x_off"][i] + meta[i, 3] * meta[i, 5]].copy()
                    for i in sel])
        info = i32(np.full(len(sel), 99))
        hip.dtrsv_vbatched(int(form[0]), int(form[1]), int(form[2]), i32(ns), A.ptrs, i32(ldas), X.ptrs, i32(incs),
                           info, max(ns), handle=h)
        torch.cuda.synchronize()
        assert np.all(info.cpu().numpy() == 0)
        gx = X.all()
        for j, i in enumerate(sel):
            xo = g["trsv_x_off"][i]
            assert_bits_equal(gx[j], g["trsv_x_out"][xo:xo + meta[i, 3] * meta[i, 5]], f"trsv {form} n={ns[j]}")
    h.close()


@pytest.mark.parametrize("p", ["dgemm", "zgemm"])
def test_golden_gemm(p):
    hip = hipmod()
    g = load("gemm.npz")
    meta = g[f"{p}_meta"]
    cplx = p == "zgemm"
    dt = torch.complex128 if cplx else torch.float64
    ta, tb, m, n, k, lda, ldb, ldc = (meta[:, j] for j in range(8))
    ca = np.where(np.isin(ta, (111, 114)), k, m)
    cb = np.where(np.isin(tb, (111, 114)), n, k)
    A = Packed([g[f"{p}_a"][o:o + l * c] for o, l, c in zip(g[f"{p}_a_off"], lda, ca)], dtype=dt)
    B = Packed([g[f"{p}_b"][o:o + l * c] for o, l, c in zip(g[f"{p}_b_off"], ldb, cb)], dtype=dt)
    C = Packed([g[f"{p}_c_in"][o:o + l * c].copy() for o, l, c in zip(g[f"{p}_c_off"], ldc, n)], dtype=dt)
    ab = g[f"{p}_alpha_beta"]
    al = f64(ab[:, :2].reshape(-1)) if cplx else f64(ab[:, 0])
    be = f64(ab[:, 2:].reshape(-1)) if cplx else f64(ab[:, 2])
    info = i32(np.full(len(meta), 99))
    h = exact_handle()
    hip.gemm_vbatched(i32(ta), i32(tb), i32(m), i32(n), i32(k), al, A.ptrs, i32(lda), B.ptrs, i32(ldb), be, C.ptrs,
                      i32(ldc), info, int(m.max()), int(n.max()), int(k.max()), complex_=cplx, handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.all(info.cpu().numpy() == 0)
    gc = C.all()
    for i in range(len(meta)):
        o = g[f"{p}_c_off"][i]
        assert_bits_equal(gc[i], g[f"{p}_c_out"][o:o + ldc[i] * n[i]], f"{p} {i} {tuple(meta[i])}")


def test_golden_gemv():
    """Blas.MulMatVec (double), NoTrans and Trans in one batch, with strides
    and NaN padding rows: y bit-identical, elements between strides kept."""
    hip = hipmod()
    g = load("gemv.npz")
    meta = g["gemv_meta"]
    tr, m, n, lda, incx, incy = (meta[:, j] for j in range(6))
    xl = np.where(tr == 112, m, n)
    yl = np.where(tr == 112, n, m)
    A = Packed([g["gemv_a"][o:o + l * c] for o, l, c in zip(g["gemv_a_off"], lda, n)])
    X = Packed([g["gemv_x"][o:o + l * c] for o, l, c in zip(g["gemv_x_off"], xl, incx)])
    Y = Packed([g["gemv_y_in"][o:o + l * c].copy() for o, l, c in zip(g["gemv_y_off"], yl, incy)])
    info = i32(np.full(len(meta), 99))
    h = exact_handle()
    hip.dgemv_vbatched(i32(tr), i32(m), i32(n), A.ptrs, i32(lda), X.ptrs, i32(incx), Y.ptrs, i32(incy), info,
                       int(m.max()), int(n.max()), handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.all(info.cpu().numpy() == 0)
    gy = Y.all()
    for i in range(len(meta)):
        o = g["gemv_y_off"][i]
        assert_bits_equal(gy[i], g["gemv_y_out"][o:o + yl[i] * incy[i]], f"gemv {i} {tuple(meta[i])}")


@pytest.mark.parametrize("small", [True, False])
def test_golden_getri(small):
    """Factorization.LuInverse (double) from the stored factors and pivots:
    the inverse bit-identical, NaN padding rows untouched.  small: the
    orders up to 64 alone (one workgroup per matrix); otherwise all orders
    (up to 100: one LU solve per column)."""
    hip = hipmod()
    g = load("getri.npz")
    sel = [i for i, (n, _) in enumerate(g["getri_meta"]) if n <= 64 or not small]
    meta = g["getri_meta"][sel]
    ns, ldas = meta[:, 0], meta[:, 1]
    offs, poffs = g["getri_off"][sel], g["getri_piv_off"][sel]
    A = Packed([g["getri_lu"][o:o + l * n].copy() for o, n, l in zip(offs, ns, ldas)])
    P = Packed([g["getri_piv"][o:o + n] for o, n in zip(poffs, ns)], dtype=torch.int32)
    info = i32(np.full(len(meta), 99))
    h = exact_handle()
    hip.dgetri_vbatched(i32(ns), A.ptrs, i32(ldas), P.ptrs, info, int(ns.max()), handle=h)
    torch.cuda.synchronize()
    h.close()
    assert np.all(info.cpu().numpy() == 0)
    ga = A.all()
    for i in range(len(meta)):
        o = offs[i]
        assert_bits_equal(ga[i], g["getri_inv"][o:o + ldas[i] * ns[i]], f"getri {i} {tuple(meta[i])}")
