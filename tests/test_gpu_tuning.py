"""The autotuner's table and the launch shapes it chooses between: a handle
applies the committed table; every exact-Cholesky workgroup shape gives the
same bits (each element is the same MatFlat-order dot product, whoever
computes it); invalid shapes are rejected."""
import numpy as np
import pytest
import torch

import oracle
from gpu_util import Packed, assert_bits_equal, i32
from helpers import spd_upper_buf

pytestmark = pytest.mark.gpu


def test_handle_applies_committed_table():
    from matflat_amd import autotune, hip
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    table = autotune.load(arch)
    assert table is not None, f"no tuning table for {arch}"
    h = hip.Handle()
    for rb, w in table["fast_waves"].items():
        assert h.fast_waves(int(rb)) == w
    assert h.bucket_streams() == table["bucket_streams"]
    # the built-in defaults hold the same choices
    raw = hip.Handle(tuned=False)
    for rb, w in table["fast_waves"].items():
        assert raw.fast_waves(int(rb)) == w, rb
    assert raw.bucket_streams() == table["bucket_streams"]


@pytest.mark.parametrize("bucket", [0, 1, 2, 3, 4])
def test_every_chol_shape_gives_the_same_bits(bucket):
    from matflat_amd import autotune, hip
    rng = np.random.default_rng(bucket)
    lo, hi = max(1, 16 * bucket), 16 * (bucket + 1) - 1
    ns = [int(v) for v in rng.integers(lo, hi + 1, size=40)]
    ldas = [n + 1 for n in ns]
    bufs = [spd_upper_buf(500 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    xs = [rng.standard_normal(n) for n in ns]
    wants = []
    for i, n in enumerate(ns):
        a, x = bufs[i].copy(), xs[i].copy()
        assert oracle.dposv(n, a, ldas[i], x) == 0
        wants.append((a, x))
    h = hip.Handle(tuned=False)
    h.bind_stream()
    for g, s in autotune.chol_shapes_for(bucket):
        h.set_chol_shape(bucket, g, s)
        A, X = Packed(bufs), Packed(xs)
        info = i32(np.full(len(ns), 99))
        hip.dposv_vbatched(i32(ns), A.ptrs, i32(ldas), X.ptrs, info, hi, handle=h)
        torch.cuda.synchronize()
        assert np.all(info.cpu().numpy() == 0)
        for i, (ga, gx) in enumerate(zip(A.all(), X.all())):
            assert_bits_equal(ga, wants[i][0], f"shape {g}x{s} factor {i}")
            assert_bits_equal(gx, wants[i][1], f"shape {g}x{s} x {i}")


def test_invalid_chol_shapes_rejected():
    from matflat_amd import hip
    h = hip.Handle(tuned=False)
    with pytest.raises(hip.ArgumentError):
        h.set_chol_shape(7, 8, 2)   # 16 lanes of rows for a 128-row bucket
    with pytest.raises(hip.ArgumentError):
        h.set_chol_shape(0, 12, 2)  # not a built shape
    with pytest.raises(hip.ArgumentError):
        h.set_chol_shape(9, 64, 2)  # no such bucket
    h.set_chol_shape(3, 0, 0)
