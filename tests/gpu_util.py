"""Helpers for GPU parity tests: pack host buffers into one device buffer and
build the per-matrix device arrays the C ABI takes."""
from __future__ import annotations

import numpy as np
import torch


def bits(a: np.ndarray) -> np.ndarray:
    """Bit patterns for exact comparison.  Every NaN maps to one pattern: the
    sign and payload of a NaN produced by arithmetic differ between x86 (the
    oracle, and MatFlat's CPU path) and gfx950 and carry no meaning."""
    a = np.array(a, copy=True)
    if a.dtype == np.complex128:
        a = a.view(np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def assert_bits_equal(got: np.ndarray, want: np.ndarray, what: str = "") -> None:
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        idx = np.flatnonzero(g != w)
        gi = got.reshape(-1)[idx[:5]] if got.dtype != np.complex128 else got.view(np.float64)[idx[:5]]
        wi = want.reshape(-1)[idx[:5]] if want.dtype != np.complex128 else want.view(np.float64)[idx[:5]]
        raise AssertionError(f"{what}: {idx.size} elements differ; first at {idx[:5]}: got {gi} want {wi}")


class Packed:
    """Host buffers packed back to back in one device tensor."""

    def __init__(self, bufs: list[np.ndarray], dtype=torch.float64, align: int = 1):
        self.sizes = [int(b.size) for b in bufs]
        offs, o = [], 0
        for s in self.sizes:
            offs.append(o)
            o += (s + align - 1) // align * align
        self.offsets = offs
        self.total = max(o, 1)
        host = np.zeros(self.total, dtype=np.complex128 if dtype == torch.complex128 else
                        (np.int32 if dtype == torch.int32 else np.float64))
        for b, off in zip(bufs, offs):
            host[off:off + b.size] = b
        self.dev = torch.from_numpy(host).to("cuda")
        es = self.dev.element_size()
        self.ptrs = torch.tensor([self.dev.data_ptr() + off * es for off in offs], dtype=torch.int64,
                                 device="cuda")

    def get(self, i: int) -> np.ndarray:
        h = self.dev.cpu().numpy()
        return h[self.offsets[i]:self.offsets[i] + self.sizes[i]].copy()

    def all(self) -> list[np.ndarray]:
        h = self.dev.cpu().numpy()
        return [h[o:o + s].copy() for o, s in zip(self.offsets, self.sizes)]


def i32(v) -> torch.Tensor:
    return torch.tensor(np.asarray(v, dtype=np.int32), dtype=torch.int32, device="cuda")


def f64(v) -> torch.Tensor:
    return torch.tensor(np.asarray(v, dtype=np.float64), dtype=torch.float64, device="cuda")
