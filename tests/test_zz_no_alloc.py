"""mfh_dposv_vbatched_noalloc (the SPD solve on the library's static device
scratch, for callers that cannot allocate) against mfh_dposv_vbatched on the
same inputs: bit-identical factors, solutions and info, in both math modes,
on two calls in a row with the second batch smaller; a count above the
scratch limit is rejected as argument 2."""
import numpy as np
import pytest
import torch

from gpu_util import Packed, assert_bits_equal, i32
from helpers import spd_upper_buf

pytestmark = pytest.mark.gpu


def batch(seed, count, nmax):
    rng = np.random.default_rng(seed)
    ns = [int(v) for v in rng.integers(1, nmax + 1, size=count)]
    ldas = [n + int(rng.integers(0, 5)) for n in ns]
    bufs = [spd_upper_buf(seed * 1000 + i, n, l) for i, (n, l) in enumerate(zip(ns, ldas))]
    # one matrix that is not positive definite
    j = min(3, ns[1] - 1)
    bufs[1][j + j * ldas[1]] = -1.0e3
    xs = [np.random.default_rng(seed * 1000 + 500 + i).standard_normal(n) for i, n in enumerate(ns)]
    return ns, ldas, bufs, xs


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_noalloc_matches_allocating_entry(mode):
    from matflat_amd import hip
    h = hip.Handle()
    h.bind_stream()
    h.set_math_mode(hip.MATH_FAST if mode == "fast" else hip.MATH_EXACT)
    for call, (count, nmax) in enumerate([(400, 128), (57, 70)]):
        ns, ldas, bufs, xs = batch(31 + call, count, nmax)
        ref_a, ref_x = Packed(bufs), Packed(xs)
        got_a, got_x = Packed(bufs), Packed(xs)
        ref_info = i32(np.full(count, 99))
        got_info = i32(np.full(count, 77))
        hip.dposv_vbatched(i32(ns), ref_a.ptrs, i32(ldas), ref_x.ptrs, ref_info, nmax, handle=h)
        hip.dposv_vbatched_noalloc(i32(ns), got_a.ptrs, i32(ldas), got_x.ptrs, got_info, nmax, handle=h)
        torch.cuda.synchronize()
        ri, gi = ref_info.cpu().numpy(), got_info.cpu().numpy()
        assert np.array_equal(ri, gi), (call, ri[:8], gi[:8])
        assert ri[1] != 0 and np.count_nonzero(ri) == 1, ri[:8]
        ga, ra, gx, rx = got_a.all(), ref_a.all(), got_x.all(), ref_x.all()
        for i in range(count):
            assert_bits_equal(ga[i], ra[i], f"call {call} A {i}")
            assert_bits_equal(gx[i], rx[i], f"call {call} x {i}")
    h.close()


def test_noalloc_rejects_count_above_scratch():
    from matflat_amd import _lib, hip
    h = hip.Handle()
    n = torch.ones(_lib.NOALLOC_MAX_COUNT + 1, dtype=torch.int32, device="cuda")
    ptrs = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(hip.ArgumentError) as e:
        hip.dposv_vbatched_noalloc(n, ptrs, n, ptrs, n, 1, handle=h)
    assert e.value.position == 2
    h.close()
