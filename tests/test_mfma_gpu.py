"""GPU parity of the split-precision matrix-core forms (pcx_plan_set_precision):
* PCX_PREC_BF16X3: the radix-16 butterflies as v_mfma_f32_32x32x16_bf16 products of three-term
  BF16 splits;
* PCX_PREC_FP16X2: v_mfma_f32_32x32x16_f16 products of hi + lo FP16 splits, each butterfly's
  inputs scaled by a power of two into the FP16 range.
Target: relative L2 error <= 1e-6 against the FP64 oracle, in both directions, every order, for
N = 2^4 .. 2^24."""
import numpy as np
import pytest

from helpers import batch_buffer, to_device, unpack_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x2"])
@pytest.mark.parametrize("lg", list(range(4, 25)))
def test_split_forward_inverse(orc, pa, lg, precision):
    import torch

    n = 1 << lg
    batch = 3 if lg <= 16 else 1
    xs = [orc.gen_normal(0x5EED0005 + b, n) for b in range(batch)]
    buf = batch_buffer(xs, n, 2 * n)
    s = to_device(buf)
    d = torch.zeros_like(s)
    r = torch.zeros_like(s)
    unit = pa.fft_unit(n, precision=precision)
    unit.batch(d.data_ptr(), s.data_ptr(), batch, n, 2 * n, n, 2 * n)
    unit.batch(r.data_ptr(), d.data_ptr(), batch, n, 2 * n, n, 2 * n, inverse=True)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, n, 2 * n, batch)
    back = unpack_batch(r.cpu().numpy(), n, n, 2 * n, batch)
    for b in range(batch):
        x = xs[b].astype(np.complex128)
        err = orc.rel_l2(got[b], orc.fft64(x))
        assert err < 1e-6, (lg, b, err)
        assert orc.rel_l2(back[b], x) < 2e-6


@pytest.mark.parametrize("lg", [6, 12, 16])
def test_fp16x2_dynamic_range(orc, pa, lg):
    """The per-butterfly power-of-two scaling: transforms at 1e-30, 1e-5, 1 and 1e20 times a
    normal input (far outside FP16's 6e-8 .. 65504), one that mixes 1e-6 and 1e3 entries, and an
    all-zero one, each within the bound of its own norm."""
    import torch

    n = 1 << lg
    base = [orc.gen_normal(0x5EED0900 + b, n).astype(np.complex128) for b in range(6)]
    mix = base[4] * np.where(np.arange(n) % 3 == 0, 1e3, 1e-6)
    xs = [base[0] * 1e-30, base[1] * 1e-5, base[2], base[3] * 1e20, mix, np.zeros(n)]
    buf = batch_buffer([x.astype(np.complex64) for x in xs], n, 2 * n)
    s = to_device(buf)
    d = torch.zeros_like(s)
    unit = pa.fft_unit(n, precision="fp16x2")
    unit.batch(d.data_ptr(), s.data_ptr(), len(xs), n, 2 * n, n, 2 * n)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, n, 2 * n, len(xs))
    for b, x in enumerate(xs):
        x32 = x.astype(np.complex64).astype(np.complex128)
        if b == len(xs) - 1:
            assert np.all(got[b] == 0)
            continue
        err = orc.rel_l2(got[b], orc.fft64(x32))
        assert err < 1e-6, (lg, b, err)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x2"])
def test_split_rejects_unsupported(pa, precision):
    from packed_amd import PcxError

    with pytest.raises(PcxError):
        pa.fft_unit(1024, dtype="float64", precision=precision)
    with pytest.raises(PcxError):
        pa.fft_unit(8, precision=precision)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x2"])
@pytest.mark.parametrize("order", ["bit_reversed", "unordered"])
@pytest.mark.parametrize("lg", [4, 5, 7, 9, 10, 12, 13, 14, 15, 16, 17, 20, 21, 24])
def test_split_permuted_orders(orc, pa, lg, order, precision):
    """bit_reversed / unordered output (forward) and input (inverse) of the split forms, index for
    index against the oracle's order maps (fft.hpp:57-68, avx2_fft.hpp:188-214), within the same
    1e-6 relative L2 of FP64 as the natural order."""
    import torch

    if order == "unordered" and lg < 5:
        pytest.skip("unordered needs the 32-point finisher (N >= 32)")
    n = 1 << lg
    batch = 3 if lg <= 16 else 1
    spk = 8  # pack 8 in and out (an unordered spectrum keeps pack max(P, 8) = 8)
    dist = orc.real_size(n, 8)
    xs = [orc.gen_normal(0x5EED0920 + b, n) for b in range(batch)]
    s = to_device(batch_buffer(xs, 8, dist))
    d = torch.zeros_like(s)
    r = torch.zeros_like(s)
    unit = pa.fft_unit(n, order=order, precision=precision)
    unit.batch(d.data_ptr(), s.data_ptr(), batch, spk, dist, 8, dist)
    unit.batch(r.data_ptr(), d.data_ptr(), batch, 8, dist, spk, dist, inverse=True)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, spk, dist, batch)
    back = unpack_batch(r.cpu().numpy(), n, 8, dist, batch)
    dnp = d.cpu().numpy()
    for b in range(batch):
        x = xs[b].astype(np.complex128)
        _, want = orc.fft_unit(order, False, True, n, orc.to_packed(xs[b], 8), 8, 8, n)
        err = orc.rel_l2(got[b], want)
        assert err < 1e-6, (lg, order, b, err)
        # the inverse against the oracle's inverse of the same permuted spectrum, and the round trip
        spec = dnp[b * dist:(b + 1) * dist]
        _, want_inv = orc.fft_unit(order, True, True, n, spec, spk, 8)
        assert orc.rel_l2(back[b], want_inv) < 1e-6, (lg, order, b)
        assert orc.rel_l2(back[b], x) < 2e-6


@pytest.mark.parametrize("lg,scale", [(4, 1e36), (6, 1e35), (4, 1e-37), (8, 1e-37)])
def test_fp16x2_exponent_extremes(orc, pa, lg, scale):
    """Inputs near the ends of the FP32 exponent range: the per-butterfly scale 2^e reaches
    e = +-126 (both 2^e and 2^-e stay normal), so neither FP16 overflow nor FP16 underflow sets in."""
    import torch

    n = 1 << lg
    x = orc.gen_normal(0x5EED0910 + lg, n).astype(np.complex128) * scale
    x32 = x.astype(np.complex64)
    s = to_device(batch_buffer([x32], n, 2 * n))
    d = torch.zeros_like(s)
    pa.fft_unit(n, precision="fp16x2").batch(d.data_ptr(), s.data_ptr(), 1, n, 2 * n, n, 2 * n)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, n, 2 * n, 1)[0]
    assert np.all(np.isfinite(got))
    err = orc.rel_l2(got, orc.fft64(x32.astype(np.complex128)))
    assert err < 1e-6, err
