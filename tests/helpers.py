"""Shared helpers of the GPU parity tests: inputs from the oracle's seeded generator, device
buffers in the packed layout, and the FP32 tolerance of the FFT path."""
from __future__ import annotations

import math

import numpy as np

EPS32 = 2.0 ** -23  # FLT_EPSILON


def fft_tol(n: int) -> float:
    """Relative L2 bound for an FP32 FFT of size n: 2 * log2(n) * FLT_EPSILON."""
    return 2.0 * max(1, int(math.log2(n))) * EPS32


def to_device(buf: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(buf, np.float32)).cuda()


def batch_buffer(vectors, pack: int, dist: int) -> np.ndarray:
    """Concatenate packed vectors at a distance of `dist` floats."""
    from oracle import oracle as O

    out = np.zeros(dist * len(vectors), np.float32)
    for b, z in enumerate(vectors):
        pb = O.to_packed(z, pack)
        out[b * dist: b * dist + len(pb)] = pb
    return out


def unpack_batch(buf: np.ndarray, n: int, pack: int, dist: int, batch: int):
    from oracle import oracle as O

    return [O.from_packed(buf[b * dist:], n, pack) for b in range(batch)]
