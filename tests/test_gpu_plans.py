"""hipGraph plans of every batched entry point (mfh_plan_create_*): each
plan is replayed several times, with eager work and restored inputs between
the replays, and every replay must equal the eager call bit for bit (and the
oracle, for the MatFlat-order entry points)."""
import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, i32
from helpers import rand_buf, spd_upper_buf

pytestmark = pytest.mark.gpu


def hipmod():
    from matflat_amd import hip
    return hip


def handle(math="exact"):
    hip = hipmod()
    h = hip.Handle()
    h.set_math_mode(hip.MATH_EXACT if math == "exact" else hip.MATH_FAST)
    h.bind_stream()
    return h


def replay(plan, restore, check, times=3):
    """restore inputs, replay, check -- with unrelated eager work between."""
    junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
    for r in range(times):
        restore()
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        check(r)
        junk.mul_(1.5).sum().item()


def test_plan_dgetrf_variable_sizes_with_large():
    hip = hipmod()
    rng = np.random.default_rng(1)
    cases = [(int(rng.integers(1, 129)), int(rng.integers(1, 129))) for _ in range(60)] + [(200, 180), (300, 300)]
    cases = [(m, n, m + 1) for m, n in cases]
    bufs = [rand_buf(i, m, n, lda, dist="normal") for i, (m, n, lda) in enumerate(cases)]
    want = [b.copy() for b in bufs]
    wp = []
    for i, (m, n, lda) in enumerate(cases):
        p, s = oracle.dgetrf(m, n, want[i], lda)
        wp.append((p, s))
    A = Packed(bufs)
    P = Packed([np.zeros(m, dtype=np.int32) for m, _, _ in cases], dtype=torch.int32)
    sign = i32(np.zeros(len(cases)))
    info = i32(np.full(len(cases), 99))
    ms, ns, ldas = i32([c[0] for c in cases]), i32([c[1] for c in cases]), i32([c[2] for c in cases])
    src = A.dev.clone()
    h = handle()
    plan = hip.plan_dgetrf(ms, ns, A.ptrs, ldas, P.ptrs, sign, info, 300, 300, handle=h)
    assert plan.op() == hip._lib.OP_DGETRF

    def restore():
        A.dev.copy_(src)
        P.dev.zero_()
        info.fill_(99)

    def check(r):
        assert np.all(info.cpu().numpy() == 0)
        ga, gp, gs = A.all(), P.all(), sign.cpu().numpy()
        for i in range(len(cases)):
            assert_bits_equal(ga[i], want[i], f"replay {r} lu {i}")
            assert np.array_equal(gp[i], wp[i][0]) and gs[i] == wp[i][1]

    replay(plan, restore, check)
    plan.close()


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_plan_dpotrf_with_large_orders(math):
    """max_n 200 runs the generic kernel on the plan's own HBM scratch."""
    hip = hipmod()
    rng = np.random.default_rng(2)
    ns = [int(v) for v in rng.integers(1, 129, size=40)] + [150, 200]
    ldas = [n + 2 for n in ns]
    bufs = [spd_upper_buf(10 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    A = Packed(bufs)
    info = i32(np.full(len(ns), 99))
    src = A.dev.clone()
    h = handle(math)
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=h)
    torch.cuda.synchronize()
    eager = A.all()
    plan = hip.plan_dpotrf(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=h)
    # eager work on the same handle between capture and replay regrows nothing the plan uses
    A.dev.copy_(src)
    hip.dpotrf_vbatched(i32(ns), A.ptrs, i32(ldas), info, max(ns), handle=h)

    def restore():
        A.dev.copy_(src)
        info.fill_(99)

    def check(r):
        assert np.all(info.cpu().numpy() == 0)
        for i, got in enumerate(A.all()):
            assert_bits_equal(got, eager[i], f"replay {r} potrf {i}")
            if math == "exact" or ns[i] > 128:
                want = bufs[i].copy()
                assert oracle.dpotrf(ns[i], want, ldas[i]) == 0
                assert_bits_equal(got, want, f"replay {r} potrf oracle {i}")

    replay(plan, restore, check)


def test_plan_dtrsv():
    hip = hipmod()
    rng = np.random.default_rng(3)
    ns = [int(v) for v in rng.integers(1, 129, size=30)] + [300]
    ldas = [n + 1 for n in ns]
    abufs, xbufs = [], []
    for i, (n, lda) in enumerate(zip(ns, ldas)):
        a = rng.uniform(-1, 1, lda * n)
        for j in range(n):
            a[j * lda + j] = 2.0 + np.sqrt(n)
        abufs.append(a)
        xbufs.append(rng.uniform(-1, 1, n))
    A, X = Packed(abufs), Packed(xbufs)
    info = i32(np.full(len(ns), 99))
    src = X.dev.clone()
    plan = hip.plan_dtrsv(122, 112, 131, i32(ns), A.ptrs, i32(ldas), X.ptrs, i32([1] * len(ns)), info, max(ns),
                          handle=handle())
    wants = []
    for i in range(len(ns)):
        w = xbufs[i].copy()
        oracle.dtrsv(122, 112, ns[i], abufs[i], ldas[i], w, 1)
        wants.append(w)

    def restore():
        X.dev.copy_(src)

    def check(r):
        assert np.all(info.cpu().numpy() == 0)
        for i, got in enumerate(X.all()):
            assert_bits_equal(got, wants[i], f"replay {r} trsv {i}")

    replay(plan, restore, check)


@pytest.mark.parametrize("mixed", [False, True])
def test_plan_dgesv(mixed):
    """The blocked LU solvers: every replay equals the eager call bit for bit
    (x, berr, iters, info); the idle refinement steps end at once."""
    hip = hipmod()
    rng = np.random.default_rng(4)
    ns = [128, 256, 64, 300, 17, 256]
    bufs, bs = [], []
    for i, n in enumerate(ns):
        a = rng.standard_normal((n, n)) + n * np.eye(n) * (0.5 if i % 2 else 0.05)
        bufs.append(np.asfortranarray(a).reshape(-1, order="F"))
        bs.append(rng.standard_normal(n))
    A, B = Packed(bufs), Packed(bs)
    X = Packed([np.zeros(n) for n in ns])
    info = i32(np.full(len(ns), 99))
    iters = i32(np.full(len(ns), -9))
    berr = torch.full((len(ns),), -1.0, dtype=torch.float64, device="cuda")
    h = handle()
    n_t, lda_t = i32(ns), i32(ns)
    if mixed:
        hip.dgesv_mixed_vbatched(n_t, A.ptrs, lda_t, B.ptrs, X.ptrs, 10, 1e-15, iters, berr, info, 300, handle=h)
    else:
        hip.dgesv_vbatched(n_t, A.ptrs, lda_t, B.ptrs, X.ptrs, berr, info, 300, handle=h)
    torch.cuda.synchronize()
    ex, eb, ei, eit = X.dev.clone(), berr.clone(), info.clone(), iters.clone()
    assert torch.all(ei == 0)
    if mixed:
        plan = hip.plan_dgesv_mixed(n_t, A.ptrs, lda_t, B.ptrs, X.ptrs, 10, 1e-15, iters, berr, info, 300, handle=h)
    else:
        plan = hip.plan_dgesv(n_t, A.ptrs, lda_t, B.ptrs, X.ptrs, berr, info, 300, handle=h)

    def restore():
        X.dev.zero_()
        berr.fill_(-1.0)
        info.fill_(99)
        iters.fill_(-9)

    def check(r):
        assert torch.equal(X.dev, ex), r
        assert torch.equal(berr, eb) and torch.equal(info, ei), r
        if mixed:
            assert torch.equal(iters, eit), r

    replay(plan, restore, check)
