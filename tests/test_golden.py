"""The oracle against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py): a change to the C restatement of
MatFlat that alters any output bit fails here.  The fixtures' shapes follow
the reference's tests (LuTests.cs:103-122, CholeskyTests.cs,
BlasTests.SolveTriangular*.cs, BlasTests.MulMatMat*.cs,
BlasTests.MulMatVecDouble.cs, LuInverseTests.cs)."""
import os

import numpy as np

import oracle
from gpu_util import assert_bits_equal

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_lu_fixtures():
    g = load("lu.npz")
    for i, (m, n, lda, sign) in enumerate(g["lu_meta"]):
        off, poff = g["lu_off"][i], g["lu_piv_off"][i]
        a = g["lu_in"][off:off + lda * n].copy()
        piv, s = oracle.dgetrf(int(m), int(n), a, int(lda))
        assert_bits_equal(a, g["lu_out"][off:off + lda * n], f"lu {i}")
        assert np.array_equal(piv, g["lu_piv"][poff:poff + m]) and s == sign


def test_cholesky_and_posv_fixtures():
    g = load("cholesky_posv.npz")
    for i, (n, lda, info) in enumerate(g["chol_meta"]):
        off, boff = g["chol_off"][i], g["posv_off"][i]
        a = g["chol_in"][off:off + lda * n].copy()
        assert oracle.dpotrf(int(n), a, int(lda)) == info
        assert_bits_equal(a, g["chol_out"][off:off + lda * n], f"chol {i}")
        a = g["chol_in"][off:off + lda * n].copy()
        x = g["posv_b"][boff:boff + n].copy()
        assert oracle.dposv(int(n), a, int(lda), x) == 0
        assert_bits_equal(x, g["posv_x"][boff:boff + n], f"posv {i}")


def test_trsv_fixtures():
    g = load("trsv.npz")
    for i, (uplo, trans, diag, n, lda, inc) in enumerate(g["trsv_meta"]):
        off, xoff = g["trsv_off"][i], g["trsv_x_off"][i]
        a = g["trsv_a"][off:off + lda * n].copy()
        if diag == 132:
            for j in range(n):
                a[j * lda + j] = 1.0
        x = g["trsv_x_in"][xoff:xoff + n * inc].copy()
        oracle.dtrsv(int(uplo), int(trans), int(n), a, int(lda), x, int(inc))
        assert_bits_equal(x, g["trsv_x_out"][xoff:xoff + n * inc], f"trsv {i}")


def test_gemm_fixtures():
    g = load("gemm.npz")
    for p, fn in (("dgemm", oracle.dgemm), ("zgemm", oracle.zgemm)):
        for i, (ta, tb, m, n, k, lda, ldb, ldc) in enumerate(g[f"{p}_meta"]):
            ar, ai, br, bi = g[f"{p}_alpha_beta"][i]
            alpha, beta = (complex(ar, ai), complex(br, bi)) if p == "zgemm" else (ar, br)
            ca = k if ta in (111, 114) else m
            cb = n if tb in (111, 114) else k
            a0, b0, c0 = g[f"{p}_a_off"][i], g[f"{p}_b_off"][i], g[f"{p}_c_off"][i]
            a = g[f"{p}_a"][a0:a0 + lda * ca]
            b = g[f"{p}_b"][b0:b0 + ldb * cb]
            c = g[f"{p}_c_in"][c0:c0 + ldc * n].copy()
            fn(int(ta), int(tb), int(m), int(n), int(k), alpha, a, int(lda), b, int(ldb), beta, c, int(ldc))
            assert_bits_equal(c, g[f"{p}_c_out"][c0:c0 + ldc * n], f"{p} {i}")


def test_gemv_fixtures():
    g = load("gemv.npz")
    for i, (trans, m, n, lda, incx, incy) in enumerate(g["gemv_meta"]):
        xl, yl = (m, n) if trans == 112 else (n, m)
        a0, x0, y0 = g["gemv_a_off"][i], g["gemv_x_off"][i], g["gemv_y_off"][i]
        a = g["gemv_a"][a0:a0 + lda * n]
        x = g["gemv_x"][x0:x0 + xl * incx]
        y = g["gemv_y_in"][y0:y0 + yl * incy].copy()
        assert oracle.dgemv(int(trans), int(m), int(n), a, int(lda), x, int(incx), y, int(incy)) == 0
        assert_bits_equal(y, g["gemv_y_out"][y0:y0 + yl * incy], f"gemv {i}")


def test_getri_fixtures():
    g = load("getri.npz")
    for i, (n, lda) in enumerate(g["getri_meta"]):
        o, po = g["getri_off"][i], g["getri_piv_off"][i]
        a = g["getri_lu"][o:o + lda * n].copy()
        assert oracle.dgetri(int(n), a, int(lda), g["getri_piv"][po:po + n]) == 0
        assert_bits_equal(a, g["getri_inv"][o:o + lda * n], f"getri {i}")


def test_complex_fixtures():
    g = load("complex.npz")
    for i, (m, n, lda, sign) in enumerate(g["zlu_meta"]):
        off, poff = g["zlu_off"][i], g["zlu_piv_off"][i]
        a = g["zlu_in"][off:off + lda * n].copy()
        piv, s = oracle.zgetrf(int(m), int(n), a, int(lda))
        assert_bits_equal(a, g["zlu_out"][off:off + lda * n], f"zlu {i}")
        assert np.array_equal(piv, g["zlu_piv"][poff:poff + m]) and s == sign
    infos = []
    for i, (n, lda, info) in enumerate(g["zchol_meta"]):
        off = g["zchol_off"][i]
        a = g["zchol_in"][off:off + lda * n].copy()
        assert oracle.zpotrf(int(n), a, int(lda)) == info
        assert_bits_equal(a, g["zchol_out"][off:off + lda * n], f"zchol {i}")
        infos.append(int(info))
    assert infos[-1] == 6 and all(v == 0 for v in infos[:-1])
    for i, (uplo, trans, n, lda, inc) in enumerate(g["ztrsv_meta"]):
        off, xoff = g["ztrsv_off"][i], g["ztrsv_x_off"][i]
        a = g["ztrsv_a"][off:off + lda * n].copy()
        x = g["ztrsv_x_in"][xoff:xoff + n * inc].copy()
        oracle.ztrsv(int(uplo), int(trans), int(n), a, int(lda), x, int(inc))
        assert_bits_equal(x, g["ztrsv_x_out"][xoff:xoff + n * inc], f"ztrsv {i}")


def test_gevd_fixtures():
    g = load("gevd.npz")
    for i, (n, lda, ldb, info) in enumerate(g["gevd_meta"]):
        ao, bo, wo = g["gevd_a_off"][i], g["gevd_b_off"][i], g["gevd_w_off"][i]
        a = g["gevd_a_in"][ao:ao + lda * n].copy()
        b = g["gevd_b_in"][bo:bo + ldb * n].copy()
        w = np.zeros(n)
        assert oracle.dgevd(int(n), a, int(lda), b, int(ldb), w) == info
        assert_bits_equal(a, g["gevd_a_out"][ao:ao + lda * n], f"gevd a {i}")
        assert_bits_equal(b, g["gevd_b_out"][bo:bo + ldb * n], f"gevd b {i}")
        assert_bits_equal(w, g["gevd_w"][wo:wo + n], f"gevd w {i}")
