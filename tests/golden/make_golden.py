"""Writes the golden fixtures tests/golden/*.npz from the oracle (the C
restatement of MatFlat in oracle/matflat_oracle.c).

MatFlat's own tests hold no golden vectors for this path (they compare with
OpenBLAS at 1e-12), so these fixtures are the build's own: seeded inputs and
the oracle's bit-exact outputs for the shapes the reference's tests use
(LuTests.cs:103-122, CholeskyTests.cs, BlasTests.SolveTriangular*.cs,
BlasTests.MulMatMat*.cs, BlasTests.MulMatVecDouble.cs, LuInverseTests.cs,
GevdTests.cs; the complex LU, Cholesky and triangular solve on the same
shapes), stored with
allow_pickle=False arrays only.  The
GPU tests compare the HIP path against them without running the oracle.

Run from the repo root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

LU_SHAPES = [(1, 1, 1), (1, 1, 3), (2, 2, 2), (2, 2, 3), (3, 3, 3), (3, 3, 5), (3, 1, 3), (3, 1, 5), (1, 3, 1),
             (1, 3, 3), (3, 4, 3), (3, 4, 5), (4, 3, 4), (4, 3, 5), (11, 23, 11), (11, 23, 17), (16, 8, 16),
             (16, 8, 32), (23, 11, 23), (23, 11, 31), (64, 64, 64)]  # (m, n, lda): LuTests.cs:103-122, plus n = 64
CHOL_N = [1, 3, 11, 16, 23, 33, 100]
TRSV_N = [1, 3, 11, 16, 33]
# (m, n, lda, incx, incy): BlasTests.MulMatVecDouble.cs:9-22 (NoTrans) and
# :56-63 (Trans), plus shapes with more outputs than a wavefront's 64 lanes
GEMV_NOTRANS = [(1, 1, 1, 1, 1), (1, 1, 3, 2, 4), (2, 2, 2, 1, 1), (2, 2, 4, 3, 3), (3, 3, 3, 1, 1),
                (3, 3, 5, 6, 7), (1, 5, 1, 1, 1), (1, 5, 3, 2, 4), (4, 1, 4, 1, 1), (4, 1, 6, 5, 7),
                (4, 9, 4, 1, 1), (4, 9, 5, 3, 2), (8, 3, 8, 1, 1), (8, 3, 9, 2, 2)]
GEMV_TRANS = [(1, 1, 1, 1, 1), (1, 1, 3, 2, 4), (2, 2, 2, 1, 1), (2, 2, 3, 2, 3), (3, 3, 3, 1, 1),
              (3, 3, 4, 3, 2), (1, 5, 1, 1, 1), (1, 5, 2, 2, 2), (4, 1, 4, 1, 1), (4, 1, 5, 3, 4),
              (4, 9, 4, 1, 1), (4, 9, 5, 3, 2), (8, 3, 8, 1, 1), (8, 3, 9, 2, 2)]
GEMV_LARGE = [(64, 100, 67, 3, 2), (130, 40, 133, 1, 1), (70, 129, 71, 2, 1)]
GETRI_SHAPES = [(1, 1), (1, 3), (2, 2), (2, 3), (3, 3), (3, 5), (5, 5), (5, 8), (10, 10), (10, 17), (64, 64),
                (100, 103)]  # (n, lda): LuInverseTests.cs:75-84, plus n = 64 and 100
GEVD_SHAPES = [(1, 1, 1), (1, 3, 2), (2, 2, 2), (2, 4, 3), (3, 3, 3), (3, 4, 5), (5, 5, 5), (5, 7, 8), (10, 10, 10),
               (10, 17, 11), (16, 19, 16), (64, 64, 64), (100, 101, 100), (128, 128, 130)]  # GevdTests.cs:94-105, larger
GEMM_SHAPES = [(1, 1, 1), (3, 4, 5), (8, 16, 4), (17, 9, 13), (19, 21, 24)]


def pack(bufs):
    off = np.concatenate([[0], np.cumsum([b.size for b in bufs])[:-1]]).astype(np.int64)
    return np.concatenate(bufs) if bufs else np.zeros(0), off


def lu_fixture(rng):
    meta, ins, outs, pivs = [], [], [], []
    for m, n, lda in LU_SHAPES:
        a = np.full(lda * n, np.nan)
        for j in range(n):
            a[j * lda:j * lda + m] = rng.standard_normal(m)
        lu = a.copy()
        piv, sign = oracle.dgetrf(m, n, lu, lda)
        meta.append((m, n, lda, sign))
        ins.append(a)
        outs.append(lu)
        pivs.append(piv.astype(np.int32))
    a_in, off = pack(ins)
    a_out, _ = pack(outs)
    piv, poff = pack(pivs)
    return {"lu_meta": np.array(meta, dtype=np.int64), "lu_off": off, "lu_in": a_in, "lu_out": a_out,
            "lu_piv": piv.astype(np.int32), "lu_piv_off": poff}


def chol_fixture(rng):
    meta, ins, outs, bs, xs = [], [], [], [], []
    for n in CHOL_N:
        lda = (n + 7) // 8 * 8
        bm = rng.standard_normal((n, n))
        spd = bm @ bm.T + n * np.eye(n)
        a = np.full(lda * n, np.nan)  # strict lower triangle NaN: never read (CholeskyTests.cs:333)
        for j in range(n):
            a[j * lda:j * lda + j + 1] = spd[:j + 1, j]
        l = a.copy()
        info = oracle.dpotrf(n, l, lda)
        b = rng.standard_normal(n)
        x = b.copy()
        lx = a.copy()
        assert oracle.dposv(n, lx, lda, x) == 0
        meta.append((n, lda, info))
        ins.append(a)
        outs.append(l)
        bs.append(b)
        xs.append(x)
    a_in, off = pack(ins)
    a_out, _ = pack(outs)
    b, boff = pack(bs)
    x, _ = pack(xs)
    return {"chol_meta": np.array(meta, dtype=np.int64), "chol_off": off, "chol_in": a_in, "chol_out": a_out,
            "posv_b": b, "posv_x": x, "posv_off": boff}


def trsv_fixture(rng):
    meta, ins, xin, xout = [], [], [], []
    for uplo in (121, 122):
        for trans in (111, 112):
            for diag in (131, 132):
                for n in TRSV_N:
                    lda, inc = n + 1, 2
                    a = rng.uniform(-1, 1, lda * n)  # non-diagonal triangles
                    for j in range(n):
                        a[j * lda + j] = 2.0 + abs(a[j * lda + j]) * np.sqrt(n)
                    x = rng.uniform(-1, 1, n * inc)
                    want = x.copy()
                    if diag == 131:
                        oracle.dtrsv(uplo, trans, n, a, lda, want, inc)
                    else:
                        au = a.copy()
                        for j in range(n):
                            au[j * lda + j] = 1.0
                        oracle.dtrsv(uplo, trans, n, au, lda, want, inc)
                    meta.append((uplo, trans, diag, n, lda, inc))
                    ins.append(a)
                    xin.append(x)
                    xout.append(want)
    a, off = pack(ins)
    x0, xoff = pack(xin)
    x1, _ = pack(xout)
    return {"trsv_meta": np.array(meta, dtype=np.int64), "trsv_off": off, "trsv_a": a, "trsv_x_off": xoff,
            "trsv_x_in": x0, "trsv_x_out": x1}


def gemm_fixture(rng, cplx):
    ops = (111, 112, 113, 114) if cplx else (111, 112)
    meta, As, Bs, Cs, Co, ab = [], [], [], [], [], []
    for m, n, k in GEMM_SHAPES:
        for ta in ops:
            for tb in ops:
                ra, ca = (m, k) if ta in (111, 114) else (k, m)
                rb, cb = (k, n) if tb in (111, 114) else (n, k)
                lda, ldb, ldc = ra + 1, rb + 2, m + 3

                def gen(size):
                    v = rng.standard_normal(size)
                    return v + 1j * rng.standard_normal(size) if cplx else v

                a, b, c = gen(lda * ca), gen(ldb * cb), gen(ldc * n)
                if cplx:
                    alpha, beta = complex(*rng.uniform(-1, 1, 2)), complex(*rng.uniform(-1, 1, 2))
                else:
                    alpha, beta = float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))
                if len(meta) % 4 == 0:
                    alpha, beta = 1.0, 0.0  # MatFlat's own C = op(A) op(B)
                out = c.copy()
                (oracle.zgemm if cplx else oracle.dgemm)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, out, ldc)
                meta.append((ta, tb, m, n, k, lda, ldb, ldc))
                As.append(a)
                Bs.append(b)
                Cs.append(c)
                Co.append(out)
                ab.append((complex(alpha).real, complex(alpha).imag, complex(beta).real, complex(beta).imag))
    p = "zgemm" if cplx else "dgemm"
    a, aoff = pack(As)
    b, boff = pack(Bs)
    c, coff = pack(Cs)
    co, _ = pack(Co)
    return {f"{p}_meta": np.array(meta, dtype=np.int64), f"{p}_a": a, f"{p}_a_off": aoff, f"{p}_b": b,
            f"{p}_b_off": boff, f"{p}_c_in": c, f"{p}_c_off": coff, f"{p}_c_out": co,
            f"{p}_alpha_beta": np.array(ab, dtype=np.float64)}


def gemv_fixture(rng):
    """Blas.MulMatVec (double): padding rows of A and the elements of x and y
    between strides hold values the routine must neither read into the
    result nor overwrite (y keeps them, as the reference's tests check)."""
    cases = [(111, *s) for s in GEMV_NOTRANS + GEMV_LARGE] + [(112, *s) for s in GEMV_TRANS + GEMV_LARGE]
    meta, As, Xs, Yi, Yo = [], [], [], [], []
    for trans, m, n, lda, incx, incy in cases:
        a = np.full(lda * n, np.nan)
        for j in range(n):
            a[j * lda:j * lda + m] = rng.standard_normal(m)
        xl, yl = (m, n) if trans == 112 else (n, m)
        x = rng.standard_normal(xl * incx)
        y = rng.standard_normal(yl * incy)
        out = y.copy()
        assert oracle.dgemv(trans, m, n, a, lda, x, incx, out, incy) == 0
        meta.append((trans, m, n, lda, incx, incy))
        As.append(a)
        Xs.append(x)
        Yi.append(y)
        Yo.append(out)
    a, aoff = pack(As)
    x, xoff = pack(Xs)
    y, yoff = pack(Yi)
    yo, _ = pack(Yo)
    return {"gemv_meta": np.array(meta, dtype=np.int64), "gemv_a": a, "gemv_a_off": aoff, "gemv_x": x,
            "gemv_x_off": xoff, "gemv_y_in": y, "gemv_y_off": yoff, "gemv_y_out": yo}


def getri_fixture(rng):
    """Factorization.LuInverse (double) applied to Factorization.Lu's output;
    padding rows hold NaN, which must survive."""
    meta, lus, pivs, outs = [], [], [], []
    for n, lda in GETRI_SHAPES:
        a = np.full(lda * n, np.nan)
        for j in range(n):
            a[j * lda:j * lda + n] = rng.standard_normal(n)
        piv, _ = oracle.dgetrf(n, n, a, lda)
        inv = a.copy()
        assert oracle.dgetri(n, inv, lda, piv) == 0
        meta.append((n, lda))
        lus.append(a)
        pivs.append(piv.astype(np.int32))
        outs.append(inv)
    lu, off = pack(lus)
    inv, _ = pack(outs)
    piv, poff = pack(pivs)
    return {"getri_meta": np.array(meta, dtype=np.int64), "getri_lu": lu, "getri_off": off,
            "getri_piv": piv.astype(np.int32), "getri_piv_off": poff, "getri_inv": inv}


def complex_fixture(rng):
    """Complex LU on the LuTests shapes, complex Cholesky (NaN lower
    triangle, one failing matrix) and the eight complex triangular-solve
    forms (Factorization.Lu.cs:352-406, Factorization.Cholesky.cs:195-249,
    Blas.SolveTriangular.cs:280-395)."""
    def crand(k):
        return rng.standard_normal(k) + 1j * rng.standard_normal(k)

    out = {}
    meta, ins, outs, pivs = [], [], [], []
    for m, n, lda in LU_SHAPES:
        a = np.full(lda * n, complex(np.nan, np.nan))
        for j in range(n):
            a[j * lda:j * lda + m] = crand(m)
        lu = a.copy()
        piv, sign = oracle.zgetrf(m, n, lu, lda)
        meta.append((m, n, lda, sign))
        ins.append(a)
        outs.append(lu)
        pivs.append(piv.astype(np.int32))
    a_in, off = pack(ins)
    a_out, _ = pack(outs)
    piv, poff = pack(pivs)
    out.update({"zlu_meta": np.array(meta, dtype=np.int64), "zlu_off": off, "zlu_in": a_in, "zlu_out": a_out,
                "zlu_piv": piv.astype(np.int32), "zlu_piv_off": poff})
    meta, ins, outs = [], [], []
    for n in CHOL_N + [9]:
        lda = (n + 7) // 8 * 8
        bm = crand(n * n).reshape(n, n)
        hpd = bm @ bm.conj().T + n * np.eye(n)
        if n == 9:
            hpd[5, 5] = -1.0  # fails at row 5, after the partial writes of rows 0..5
        a = np.full(lda * n, complex(np.nan, np.nan))
        for j in range(n):
            a[j * lda:j * lda + j + 1] = hpd[:j + 1, j]
        l = a.copy()
        info = oracle.zpotrf(n, l, lda)
        meta.append((n, lda, info))
        ins.append(a)
        outs.append(l)
    a_in, off = pack(ins)
    a_out, _ = pack(outs)
    out.update({"zchol_meta": np.array(meta, dtype=np.int64), "zchol_off": off, "zchol_in": a_in,
                "zchol_out": a_out})
    meta, ins, xin, xout = [], [], [], []
    for uplo in (121, 122):
        for trans in (111, 112, 113, 114):
            for n in TRSV_N:
                lda, inc = n + 1, 2
                a = crand(lda * n)
                for j in range(n):
                    a[j * lda + j] = (2.0 + np.sqrt(n)) * np.exp(1j * rng.uniform(0, 2 * np.pi))
                x = crand(n * inc)
                want = x.copy()
                assert oracle.ztrsv(uplo, trans, n, a, lda, want, inc) == 0
                meta.append((uplo, trans, n, lda, inc))
                ins.append(a)
                xin.append(x)
                xout.append(want)
    a, off = pack(ins)
    x0, xoff = pack(xin)
    x1, _ = pack(xout)
    out.update({"ztrsv_meta": np.array(meta, dtype=np.int64), "ztrsv_off": off, "ztrsv_a": a, "ztrsv_x_off": xoff,
                "ztrsv_x_in": x0, "ztrsv_x_out": x1})
    return out


def gevd_fixture(rng):
    """Factorization.Gevd (double): A = X X^T, B = Y Y^T + I/2 with X, Y ~
    U(-1, 1) (GevdTests.cs:295-330 builds the same kind of pair), strict lower
    triangles NaN (never read), padding rows random; the last case has a B
    that is not positive definite (info = 4, partial writes in b)."""
    meta, ains, bins, aouts, bouts, ws = [], [], [], [], [], []
    cases = [(n, lda, ldb, False) for n, lda, ldb in GEVD_SHAPES] + [(6, 7, 6, True)]
    for n, lda, ldb, bad in cases:
        bufs = []
        for ld, shift in ((lda, 0.0), (ldb, 0.5)):
            buf = 2.0 * rng.random(ld * n) - 1.0
            x = buf[: ld * n].reshape(n, ld).T[:n, :n].copy()
            m = x @ x.T + shift * np.eye(n)
            v = buf[: ld * n].reshape(n, ld).T
            v[:n, :n] = m
            for j in range(n):
                v[j + 1:n, j] = np.nan
            bufs.append(buf)
        a, b = bufs
        if bad:
            b[3 + 3 * ldb] = -50.0
        w = np.zeros(n)
        ao, bo = a.copy(), b.copy()
        info = oracle.dgevd(n, ao, lda, bo, ldb, w)
        assert info == (4 if bad else 0)
        meta.append((n, lda, ldb, info))
        ains.append(a)
        bins.append(b)
        aouts.append(ao)
        bouts.append(bo)
        ws.append(w)
    a_in, aoff = pack(ains)
    b_in, boff = pack(bins)
    w, woff = pack(ws)
    return {"gevd_meta": np.array(meta, dtype=np.int64), "gevd_a_in": a_in, "gevd_a_off": aoff,
            "gevd_b_in": b_in, "gevd_b_off": boff, "gevd_a_out": pack(aouts)[0], "gevd_b_out": pack(bouts)[0],
            "gevd_w": w, "gevd_w_off": woff}


def main():
    oracle.build()
    rng = np.random.default_rng(20240610)
    np.savez_compressed(os.path.join(HERE, "lu.npz"), **lu_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "cholesky_posv.npz"), **chol_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "trsv.npz"), **trsv_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "gemm.npz"), **gemm_fixture(rng, False), **gemm_fixture(rng, True))
    np.savez_compressed(os.path.join(HERE, "gemv.npz"), **gemv_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "getri.npz"), **getri_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "complex.npz"), **complex_fixture(np.random.default_rng(20261204)))
    np.savez_compressed(os.path.join(HERE, "gevd.npz"), **gevd_fixture(np.random.default_rng(20261205)))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
