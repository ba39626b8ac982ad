"""GPU parity of the split-BF16 matrix-core form (pcx_plan_set_precision(PCX_PREC_BF16X3)):
the radix-16 butterflies run as v_mfma_f32_32x32x16_bf16 products of three-term BF16 splits.
Target: relative L2 error <= 1e-6 against the FP64 oracle, in both directions."""
import numpy as np
import pytest

from helpers import batch_buffer, to_device, unpack_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


@pytest.mark.parametrize("lg", list(range(4, 21)))
def test_bf16x3_forward_inverse(orc, pa, lg):
    import torch

    n = 1 << lg
    batch = 3 if lg <= 16 else 1
    xs = [orc.gen_normal(0x5EED0005 + b, n) for b in range(batch)]
    buf = batch_buffer(xs, n, 2 * n)
    s = to_device(buf)
    d = torch.zeros_like(s)
    r = torch.zeros_like(s)
    unit = pa.fft_unit(n, precision="bf16x3")
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


def test_bf16x3_rejects_unsupported(pa):
    from packed_amd import PcxError

    with pytest.raises(PcxError):
        pa.fft_unit(1024, order="bit_reversed", precision="bf16x3")
    with pytest.raises(PcxError):
        pa.fft_unit(8, precision="bf16x3")
