"""GPU parity of batch storages (two-axis md::dynamic_storage, mdstorage.hpp:1478-1499): one
fft_unit call transforms every slice; results match the oracle per slice and one call per slice
bit for bit; padding between slices stays untouched."""
import numpy as np
import pytest

from helpers import fft_tol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _fill(orc, st, seeds, n_fill=None):
    import torch

    n_fill = st.inner if n_fill is None else n_fill
    buf = np.zeros(st.row_stride * st.outer, np.dtype(st.dtype))
    xs = []
    for r in range(st.outer):
        x = orc.gen_uniform(seeds + r, n_fill).astype(np.complex128)
        idx = r * st.row_stride + orc.pidx(np.arange(n_fill), st.pack)
        buf[idx], buf[idx + st.pack] = x.real, x.imag
        xs.append(x)
    st.data.copy_(torch.from_numpy(buf))
    return xs


def _rows(orc, st):
    h = st.data.cpu().numpy()
    idx = orc.pidx(np.arange(st.inner), st.pack)
    return [h[r * st.row_stride + idx] + 1j * h[r * st.row_stride + idx + st.pack] for r in range(st.outer)], h


@pytest.mark.parametrize("lg,outer,pack,alignment", [(6, 50, 8, 8), (12, 19, 1, 64), (13, 7, 16, 8), (16, 5, 8, 8),
                                                     (18, 2, 4, 32)])
def test_storage_forward_and_round_trip(orc, pa, lg, outer, pack, alignment):
    import torch

    n = 1 << lg
    st = pa.md_storage(n, outer, pack, alignment)
    xs = _fill(orc, st, 0x5EED0800)
    unit = pa.fft_unit(n)
    unit(st)
    torch.cuda.synchronize()
    got, h = _rows(orc, st)
    for r in range(outer):
        assert orc.rel_l2(got[r], orc.fft64(xs[r])) < fft_tol(n)
        pad = h[r * st.row_stride + orc.real_size(n, pack): (r + 1) * st.row_stride]
        assert np.all(pad == 0)
    unit.ifft(st)
    torch.cuda.synchronize()
    back, _ = _rows(orc, st)
    for r in range(outer):
        assert orc.rel_l2(back[r], xs[r]) < 2 * fft_tol(n)


def test_storage_zero_extension_and_per_slice_equality(orc, pa):
    """A shorter, padded source storage (3000 values per slice, pack 1, alignment 64: 6016 reals per
    slice) into a 4096-point destination, against one call per slice."""
    import torch

    n, outer = 4096, 23
    src = pa.md_storage(3000, outer, 1, 64)
    assert src.row_stride == 6016
    xs = _fill(orc, src, 0x5EED0900)
    dst = pa.md_storage(n, outer)
    unit = pa.fft_unit(n)
    unit(dst, src)
    one = pa.md_storage(n, outer)
    for r in range(outer):
        unit(one[r], src[r])
    torch.cuda.synchronize()
    assert torch.equal(dst.data, one.data)
    got, _ = _rows(orc, dst)
    for r in range(outer):
        want = orc.fft64(np.concatenate([xs[r], np.zeros(n - 3000)]))
        assert orc.rel_l2(got[r], want) < fft_tol(n)


def test_storage_double_bit_reversed(orc, pa):
    import torch

    n, outer = 1 << 17, 3
    st = pa.md_storage(n, outer, dtype="float64")
    assert st.pack == 4 and st.row_stride == 2 * n
    xs = _fill(orc, st, 0x5EED0A00)
    pa.fft_unit(n, order="bit_reversed", dtype="float64")(st)
    torch.cuda.synchronize()
    got, _ = _rows(orc, st)
    perm = orc.slot_freq("bit_reversed", n)
    for r in range(outer):
        assert orc.rel_l2(got[r], orc.fft64(xs[r])[perm]) < 4 * 17 * 2.0 ** -52


def test_storage_errors(pa):
    from packed_amd import PcxError

    unit = pa.fft_unit(1024)
    with pytest.raises(PcxError):
        unit(pa.md_storage(512, 2))
    with pytest.raises(PcxError):
        unit(pa.md_storage(1024, 2), pa.md_storage(1024, 3))
    with pytest.raises(PcxError):
        unit(pa.md_storage(1024, 2, dtype="float64"))


def test_storage_fxcorr_channels(orc, pa):
    """fxcorr_unit on a batch storage: every slice a channel sharing one kernel spectrum, in place
    and into another storage, against the direct FP64 correlation per channel."""
    import torch

    L, M, ch = 20000, 300, 6
    g = orc.gen_uniform(0x5EED0D00, M)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g), 2048)
    sig = pa.md_storage(L, ch, 8, 64)
    xs = _fill(orc, sig, 0x5EED0E00)
    out = pa.md_storage(L, ch, 1, 8)
    xc(sig, out)
    xc(sig)
    torch.cuda.synchronize()
    got_in, _ = _rows(orc, sig)
    got_out, _ = _rows(orc, out)
    for c in range(ch):
        want = orc.xcorr_naive(xs[c], g)
        assert orc.rel_l2(got_out[c], want) < 4e-6
        assert orc.rel_l2(got_in[c], want) < 4e-6
