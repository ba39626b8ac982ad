"""GPU parity of the FFT across a range of vectors (pcx_fft_across; the transform of the
reference's fft_unit_par, fft.hpp:2168-2572, whose own output order is wrong -- its test is
disabled at test_fft.cpp:789 -- so the FP64 DFT along the vector index is the oracle, in natural
order)."""
import numpy as np
import pytest

from helpers import fft_tol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _run(orc, pa, n, length, sp, dp, inverse=False, normalized=True, inplace=False, dtype="float32"):
    import torch

    zs = [orc.gen_normal(0x5EED0070 + i, length) for i in range(n)]
    if dtype == "float64":  # full 53-bit significands
        zs = [z.astype(np.complex128) * (1 + 2.0 ** -30) for z in zs]
    src = [pa.vector.from_complex(z, sp, dtype=dtype) for z in zs]
    dst = src if inplace else [pa.vector(length, dp, dtype=dtype) for _ in range(n)]
    unit = pa.fft_unit(n, dtype=dtype)
    unit.across(dst, src, inverse=inverse, normalized=normalized)
    torch.cuda.synchronize()
    got = np.stack([v.to_complex() for v in dst])
    stack = np.stack([z.astype(np.complex128) for z in zs])
    want = np.fft.ifft(stack, axis=0) * (1 if normalized else n) if inverse else np.fft.fft(stack, axis=0)
    return orc.rel_l2(got.ravel(), want.ravel())


@pytest.mark.parametrize("lg", list(range(1, 14)))
def test_across_sizes(orc, pa, lg):
    n = 1 << lg
    for length in ((1, 37, 300) if lg <= 10 else (1, 19)):
        assert _run(orc, pa, n, length, 8, 8) < fft_tol(n), (n, length)


@pytest.mark.parametrize("lg", [1, 4, 7, 9, 11])
def test_across_sizes_f64(orc, pa, lg):
    """fft_unit_par<double> (fft.hpp:2168-2172 admits double): within 4 log2(N) DBL_EPSILON."""
    n = 1 << lg
    for length in (1, 33):
        err = _run(orc, pa, n, length, 4, 4, dtype="float64")
        assert err < 4 * max(1, lg) * 2.0 ** -52, (n, length, err)
    assert _run(orc, pa, n, 17, 1, 4, inverse=True, dtype="float64") < 4 * max(1, lg) * 2.0 ** -52


@pytest.mark.parametrize("packs", [(1, 1), (1, 8), (8, 1), (4, 64)])
def test_across_packs_and_inverse(orc, pa, packs):
    sp, dp = packs
    assert _run(orc, pa, 64, 513, sp, dp) < fft_tol(64)
    assert _run(orc, pa, 64, 513, sp, dp, inverse=True) < fft_tol(64)
    assert _run(orc, pa, 1024, 40, sp, dp, inverse=True, normalized=False) < fft_tol(1024)


def test_across_in_place(orc, pa):
    assert _run(orc, pa, 256, 1000, 8, 8, inplace=True) < fft_tol(256)


def test_across_errors(orc, pa):
    vs = [pa.vector(10, 8) for _ in range(8)]
    with pytest.raises(pa.PcxError):
        pa.fft_unit(8, order="bit_reversed").across(vs, vs)
    big = [pa.vector(2, 8) for _ in range(16384)]
    with pytest.raises(pa.PcxError):
        pa.fft_unit(16384).across(big, big)
    # differing packs: a destination overlapping a source is an in-place repack (fft.hpp:783)
    with pytest.raises(pa.PcxError):
        pa.fft_unit(8).across(vs_as(pa, vs, 4), vs)
    with pytest.raises(pa.PcxError):
        pa.fft_unit(16).across(vs, vs)  # 8 vectors for a 16-point transform


def vs_as(pa, vs, pack):
    """The same device memory viewed as vectors of another pack size."""
    return [pa.vector(v.size(), pack, data=v.data) for v in vs]
