"""GPU parity of the FFT unit (fft.hpp:565-915) against the FP64 oracle.

Every case: seeded inputs from the oracle's generator, the HIP path through the C ABI, and the
oracle's restatement of the same call (orc_fft_unit).  Output order is checked index-exactly
(a misplaced slot gives a relative L2 error of order 1/sqrt(N), far above the bound) and values
within 2*log2(N)*FLT_EPSILON relative L2 of FP64.
"""
import numpy as np
import pytest

from helpers import batch_buffer, fft_tol, to_device, unpack_batch

pytestmark = pytest.mark.gpu

ORDERS = ["normal", "bit_reversed", "unordered"]


def _spec_pack(order, pack):
    return max(pack, 8) if order == "unordered" else pack


def _fwd_case(O, pa, n, order, src_pack, dst_pack, batch=1, src_size=None, seed=0x5EED0001, inplace=False):
    import torch

    src_size = n if src_size is None else src_size
    dpk = _spec_pack(order, dst_pack)
    sdist = O.real_size(n, src_pack)
    ddist = O.real_size(n, dpk)
    xs = [O.gen_uniform(seed + b, n) for b in range(batch)]
    xs_cut = [x[:src_size] for x in xs]
    sbuf = batch_buffer(xs_cut, src_pack, sdist)
    unit = pa.fft_unit(n, order=order)
    s = to_device(sbuf)
    if inplace:
        assert sdist == ddist and src_pack == dpk
        d = s
    else:
        d = torch.zeros(ddist * batch, dtype=torch.float32, device="cuda")
    if batch == 1:
        unit.fft_raw(d.data_ptr(), s.data_ptr(), dst_pack, src_pack, src_size)
    else:
        unit.batch(d.data_ptr(), s.data_ptr(), batch, dst_pack, ddist, src_pack, sdist, src_size=src_size)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, dpk, ddist, batch)
    errs = []
    for b in range(batch):
        ref_src = O.to_packed(xs_cut[b], src_pack, n_alloc=n)
        _, want = O.fft_unit(order, False, True, n, ref_src, src_pack, dst_pack, src_size)
        errs.append(O.rel_l2(got[b], want))
    return max(errs)


def _inv_case(O, pa, n, order, src_pack, dst_pack, normalized=True, batch=1, seed=0x5EED0002):
    import torch

    spk = _spec_pack(order, src_pack)
    sdist = O.real_size(n, spk)
    ddist = O.real_size(n, dst_pack)
    zs = [O.gen_normal(seed + b, n) for b in range(batch)]
    sbuf = batch_buffer(zs, spk, sdist)
    unit = pa.fft_unit(n, order=order)
    s = to_device(sbuf)
    d = torch.zeros(ddist * batch, dtype=torch.float32, device="cuda")
    if batch == 1:
        unit.ifft_raw(d.data_ptr(), s.data_ptr(), normalized, dst_pack, src_pack)
    else:
        unit.batch(d.data_ptr(), s.data_ptr(), batch, dst_pack, ddist, src_pack, sdist, inverse=True,
                   normalized=normalized)
    torch.cuda.synchronize()
    got = unpack_batch(d.cpu().numpy(), n, dst_pack, ddist, batch)
    errs = []
    for b in range(batch):
        _, want = O.fft_unit(order, True, normalized, n, sbuf[b * sdist:(b + 1) * sdist], spk, dst_pack)
        errs.append(O.rel_l2(got[b], want))
    return max(errs)


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


@pytest.mark.parametrize("lg", list(range(1, 21)))
@pytest.mark.parametrize("order", ORDERS)
def test_forward_all_sizes(orc, pa, lg, order):
    n = 1 << lg
    if order == "unordered" and lg < 5:
        pytest.skip("unordered needs the 32-point finisher (N >= 32)")
    err = _fwd_case(orc, pa, n, order, 8, 8, batch=3 if lg <= 16 else 1)
    assert err < fft_tol(n), (n, order, err)


@pytest.mark.parametrize("lg", list(range(1, 21)))
@pytest.mark.parametrize("order", ORDERS)
def test_inverse_all_sizes(orc, pa, lg, order):
    n = 1 << lg
    if order == "unordered" and lg < 5:
        pytest.skip("unordered needs the 32-point finisher (N >= 32)")
    err = _inv_case(orc, pa, n, order, 8, 8, batch=2 if lg <= 16 else 1)
    assert err < fft_tol(n), (n, order, err)


@pytest.mark.parametrize("lg", [21, 22, 23, 24])
@pytest.mark.parametrize("order", ORDERS)
def test_largest_sizes(orc, pa, lg, order):
    """One transform at each four-step size above 2^20, up to the 2^24 maximum (split layout)."""
    n = 1 << lg
    assert _fwd_case(orc, pa, n, order, n, n) < fft_tol(n)
    assert _inv_case(orc, pa, n, order, n, n) < fft_tol(n)


@pytest.mark.parametrize("n", [64, 1024, 4096, 1 << 15, 1 << 16])
@pytest.mark.parametrize("packs", [(1, 1), (8, 1), (1, 8), (4, 2), ("N", "N"), ("N", 8)])
@pytest.mark.parametrize("order", ORDERS)
def test_pack_layouts(orc, pa, n, packs, order):
    sp, dp = (n if p == "N" else p for p in packs)
    assert _fwd_case(orc, pa, n, order, sp, dp) < fft_tol(n)
    assert _inv_case(orc, pa, n, order, dp, sp) < fft_tol(n)


@pytest.mark.parametrize("n,src_size", [(64, 1), (64, 37), (1024, 1000), (4096, 100), (1 << 16, 12345),
                                        (1 << 18, 1 << 17)])
@pytest.mark.parametrize("order", ORDERS)
def test_zero_extension(orc, pa, n, src_size, order):
    assert _fwd_case(orc, pa, n, order, 8, 8, src_size=src_size) < fft_tol(n)


@pytest.mark.parametrize("n", [256, 8192, 1 << 16])
@pytest.mark.parametrize("order", ORDERS)
def test_in_place(orc, pa, n, order):
    assert _fwd_case(orc, pa, n, order, 8, 8, batch=2, inplace=True) < fft_tol(n)


@pytest.mark.parametrize("n", [128, 1 << 16])
@pytest.mark.parametrize("order", ORDERS)
def test_unnormalized_inverse(orc, pa, n, order):
    assert _inv_case(orc, pa, n, order, 8, 8, normalized=False) < fft_tol(n)


@pytest.mark.parametrize("n,batch", [(64, 1000), (1024, 257), (1 << 16, 5), (1 << 20, 2)])
def test_batched(orc, pa, n, batch):
    assert _fwd_case(orc, pa, n, "normal", n, n, batch=batch) < fft_tol(n)
    assert _inv_case(orc, pa, n, "bit_reversed", n, n, batch=batch) < fft_tol(n)


@pytest.mark.parametrize("n", [1024, 1 << 16])
@pytest.mark.parametrize("order", ORDERS)
def test_round_trip(orc, pa, n, order):
    import torch

    x = orc.gen_uniform(0x5EED0001, n)
    v = pa.vector.from_complex(x, 8)
    unit = pa.fft_unit(n, order=order)
    unit(v)
    unit.ifft(v)
    torch.cuda.synchronize()
    assert orc.rel_l2(v.to_complex(), x) < 2 * fft_tol(n)


def test_reference_leaky_tone(orc, pa):
    """tests/test_fft.cpp:299-302 input exp(i*2*pi*i/N*13.37) against its closed-form DFT."""
    import torch

    for n in [64, 1024, 16384]:
        i = np.arange(n)
        f = 13.37
        x = np.exp(1j * 2 * np.pi * i / n * f).astype(np.complex64)
        k = np.arange(n)
        d = 2j * np.pi * (f - k)
        want = (1 - np.exp(d)) / (1 - np.exp(d / n))
        v = pa.vector.from_complex(x, 8)
        pa.fft_unit(n)(v)
        torch.cuda.synchronize()
        assert orc.rel_l2(v.to_complex(), want) < fft_tol(n) + 1e-6


@pytest.mark.parametrize("lg", [15, 16])
@pytest.mark.parametrize("order", ORDERS)
def test_one_pass_dif(orc, pa, lg, order):
    """N = 2^15 runs the one-pass decimation-in-frequency kernel (2^16 the four-step): batches
    that leave part of the last group of eight transforms empty, zero-extension and repacking,
    forward in every order, inverse in natural order."""
    n = 1 << lg
    for batch in (1, 9):
        assert _fwd_case(orc, pa, n, order, 8, 8, batch=batch) < fft_tol(n)
    assert _fwd_case(orc, pa, n, order, 1, n, batch=2, src_size=n - 777) < fft_tol(n)
    if order == "normal":
        assert _inv_case(orc, pa, n, order, n, n, batch=3) < fft_tol(n)
        assert _inv_case(orc, pa, n, order, 8, 2, normalized=False) < fft_tol(n)


@pytest.mark.parametrize("order", ORDERS)
def test_one_pass_dif_agrees_with_four_step(orc, pa, order):
    """The same batch through the one-pass kernel and through the two-pass four-step split."""
    import torch

    n, batch = 1 << 15, 4
    xs = [orc.gen_normal(0x5EED0040 + b, n) for b in range(batch)]
    src = to_device(batch_buffer(xs, n, 2 * n))
    outs = []
    for flag in (True, False):
        unit = pa.fft_unit(n, order=order)
        unit.set_one_pass(flag)
        d = torch.zeros_like(src)
        unit.batch(d.data_ptr(), src.data_ptr(), batch, n, 2 * n, n, 2 * n)
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy().astype(np.float64))
    diff = np.linalg.norm(outs[0] - outs[1]) / np.linalg.norm(outs[1])
    assert diff < fft_tol(n), diff


@pytest.mark.parametrize("lg", [21, 22, 23, 24])
def test_three_pass_agrees_with_four_step(orc, pa, lg):
    """FP32 natural order above 2^20 runs in three passes; the same batches through the four-step
    split agree, forward (pack 8 in, N out, zero-extended source, two chunks) and inverse
    (normalised, in place), and the forward output matches the FP64 oracle."""
    import torch

    n, batch = 1 << lg, 2
    xs = [orc.gen_normal(0x5EED0080 + b, n) for b in range(batch)]
    cut = n - 12345
    src = to_device(batch_buffer([x[:cut] for x in xs], 8, orc.real_size(n, 8)))
    sd = orc.real_size(n, 8)
    outs, backs = [], []
    for three in (True, False):
        unit = pa.fft_unit(n)
        unit.set_three_pass(three)
        unit.set_chunk(8 * n)  # one transform per chunk
        d = torch.zeros(2 * n * batch, dtype=torch.float32, device="cuda")
        unit.batch(d.data_ptr(), src.data_ptr(), batch, n, 2 * n, 8, sd, src_size=cut)
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy().astype(np.float64))
        unit.batch(d.data_ptr(), d.data_ptr(), batch, n, 2 * n, n, 2 * n, inverse=True)
        torch.cuda.synchronize()
        backs.append(d.cpu().numpy().astype(np.float64))
    diff = np.linalg.norm(outs[0] - outs[1]) / np.linalg.norm(outs[1])
    assert diff < fft_tol(n), diff
    diff = np.linalg.norm(backs[0] - backs[1]) / np.linalg.norm(backs[1])
    assert diff < fft_tol(n), diff
    for b in (0, batch - 1):
        x = np.zeros(n, dtype=np.complex128)
        x[:cut] = xs[b][:cut]
        want = orc.fft64(x)
        o = outs[0][b * 2 * n:(b + 1) * 2 * n]
        assert orc.rel_l2(o[:n] + 1j * o[n:], want) < fft_tol(n), b
        r = backs[0][b * 2 * n:(b + 1) * 2 * n]
        assert orc.rel_l2(r[:n] + 1j * r[n:], x) < 2 * fft_tol(n), b


def test_shared_plan_on_two_streams(orc, pa):
    """One four-step plan (as copies of a C++ fft_unit share it) called on two streams without a
    synchronisation between the calls: the second call waits for the first one's use of the
    plan's intermediate, so both batches come out right."""
    import torch

    n, batch = 1 << 16, 64
    unit = pa.fft_unit(n)
    xs = [[orc.gen_normal(0x5EED0070 + 100 * k + b, n) for b in range(batch)] for k in range(2)]
    srcs = [to_device(batch_buffer(x, n, 2 * n)) for x in xs]
    dsts = [torch.zeros_like(s) for s in srcs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for rep in range(3):
        for k in range(2):
            unit.batch(dsts[k].data_ptr(), srcs[k].data_ptr(), batch, n, 2 * n, n, 2 * n, stream=streams[k])
    torch.cuda.synchronize()
    for k in range(2):
        out = dsts[k].cpu().numpy()
        for b in (0, batch - 1):
            got = orc.from_packed(out[b * 2 * n:], n, n)
            assert orc.rel_l2(got, orc.fft64(xs[k][b].astype(np.complex128))) < fft_tol(n), (k, b)
