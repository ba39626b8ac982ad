"""GPU parity of the double-precision FFT unit (fft_unit<double, Order>, fft.hpp:565-915 with
T = double; default pack 4, types.hpp:70-72) against the oracle's FP64 transform.

Inputs carry a full 53-bit significand (a float32 draw plus a 2^-30-scaled second draw), so a
path that rounded anything to FP32 would miss the bound by six orders of magnitude.  Values are
checked within 4*log2(N)*DBL_EPSILON relative L2 of the oracle; output order index-exactly (a
misplaced slot costs of order 1/sqrt(N)).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS64 = 2.0 ** -52  # DBL_EPSILON


@pytest.fixture(scope="module")
def O(orc):
    return orc


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def tol64(n: int) -> float:
    return 4.0 * max(1, int(math.log2(n))) * EPS64


def x64(O, seed: int, n: int) -> np.ndarray:
    a = O.gen_uniform(seed, n).astype(np.complex128)
    b = O.gen_uniform(seed ^ 0x5A5A, n).astype(np.complex128)
    return a + b * 2.0 ** -30


def packed64(O, z, pack: int, n_alloc: int | None = None) -> np.ndarray:
    n = len(z) if n_alloc is None else n_alloc
    buf = np.zeros(O.real_size(n, pack), np.float64)
    idx = O.pidx(np.arange(len(z)), pack)
    buf[idx] = z.real
    buf[idx + pack] = z.imag
    return buf


def unpacked64(O, buf: np.ndarray, n: int, pack: int) -> np.ndarray:
    idx = O.pidx(np.arange(n), pack)
    return buf[idx] + 1j * buf[idx + pack]


def dev(buf):
    import torch

    return torch.from_numpy(np.ascontiguousarray(buf, np.float64)).cuda()


def spectrum(O, x: np.ndarray, n: int, order: str) -> np.ndarray:
    """Forward spectrum of x zero-extended to n, in the order's slot layout."""
    X = O.fft64(np.concatenate([x, np.zeros(n - len(x), np.complex128)]))
    return X if order == "normal" else X[O.slot_freq(order, n)]


ALL_LG = list(range(1, 24))


def _packs(lg):
    n = 1 << lg
    return [(4, 4), (1, 8), (n, 2)][lg % 3]


@pytest.mark.parametrize("order", ["normal", "bit_reversed"])
@pytest.mark.parametrize("lg", ALL_LG)
def test_forward_f64(O, pa, lg, order):
    import torch

    n = 1 << lg
    sp, dp = _packs(lg)
    x = x64(O, 0x5EED0101 + lg, n)
    unit = pa.fft_unit(n, order=order, dtype="float64")
    s = dev(packed64(O, x, sp))
    d = torch.zeros(O.real_size(n, dp), dtype=torch.float64, device="cuda")
    unit.fft_raw(d.data_ptr(), s.data_ptr(), dp, sp)
    torch.cuda.synchronize()
    got = unpacked64(O, d.cpu().numpy(), n, dp)
    assert O.rel_l2(got, spectrum(O, x, n, order)) < tol64(n)


@pytest.mark.parametrize("order", ["normal", "bit_reversed"])
@pytest.mark.parametrize("lg", ALL_LG)
def test_inverse_f64(O, pa, lg, order):
    import torch

    n = 1 << lg
    sp, dp = _packs(lg)[::-1]
    z = x64(O, 0x5EED0201 + lg, n)  # spectrum in slot order
    unit = pa.fft_unit(n, order=order, dtype="float64")
    s = dev(packed64(O, z, sp))
    d = torch.zeros(O.real_size(n, dp), dtype=torch.float64, device="cuda")
    unit.ifft_raw(d.data_ptr(), s.data_ptr(), True, dp, sp)
    torch.cuda.synchronize()
    nat = np.empty(n, np.complex128)
    if order == "normal":
        nat[:] = z
    else:
        nat[O.slot_freq(order, n)] = z
    want = O.fft64(nat, True) / n
    assert O.rel_l2(unpacked64(O, d.cpu().numpy(), n, dp), want) < tol64(n)


@pytest.mark.parametrize("lg", [3, 10, 13, 14, 16, 18])
def test_unnormalized_inverse_f64(O, pa, lg):
    import torch

    n = 1 << lg
    z = x64(O, 0x5EED0301, n)
    unit = pa.fft_unit(n, dtype="float64")
    v = pa.vector.from_complex(z, dtype="float64")
    unit.ifft(v, normalized=False)
    torch.cuda.synchronize()
    assert O.rel_l2(v.to_complex(), O.fft64(z, True)) < tol64(n)


@pytest.mark.parametrize("lg,src_size", [(6, 17), (12, 1000), (13, 8191), (16, 20000), (20, 1 << 19)])
def test_zero_extension_f64(O, pa, lg, src_size):
    import torch

    n = 1 << lg
    x = x64(O, 0x5EED0401, src_size)
    unit = pa.fft_unit(n, dtype="float64")
    src = pa.vector.from_complex(x, dtype="float64")
    dst = pa.vector(n, dtype="float64")
    unit(dst, src)
    torch.cuda.synchronize()
    assert O.rel_l2(dst.to_complex(), spectrum(O, x, n, "normal")) < tol64(n)


@pytest.mark.parametrize("order", ["normal", "bit_reversed"])
@pytest.mark.parametrize("lg", [5, 12, 13, 15, 17])
def test_batch_and_in_place_f64(O, pa, lg, order):
    """Three transforms at a distance beyond the footprint, in place, forward then inverse."""
    import torch

    n, batch, pack = 1 << lg, 3, 4
    dist = O.real_size(n, pack) + 64
    xs = [x64(O, 0x5EED0501 + b, n) for b in range(batch)]
    buf = np.zeros(dist * batch, np.float64)
    for b, x in enumerate(xs):
        pb = packed64(O, x, pack)
        buf[b * dist: b * dist + len(pb)] = pb
    d = dev(buf)
    unit = pa.fft_unit(n, order=order, dtype="float64")
    unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist)
    torch.cuda.synchronize()
    h = d.cpu().numpy()
    for b, x in enumerate(xs):
        assert O.rel_l2(unpacked64(O, h[b * dist:], n, pack), spectrum(O, x, n, order)) < tol64(n)
        assert np.all(h[b * dist + O.real_size(n, pack): (b + 1) * dist] == 0)  # gaps untouched
    unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist, inverse=True)
    torch.cuda.synchronize()
    h = d.cpu().numpy()
    for b, x in enumerate(xs):
        assert O.rel_l2(unpacked64(O, h[b * dist:], n, pack), x) < 2 * tol64(n)


@pytest.mark.parametrize("lg", [21, 22, 23])
def test_three_pass_f64(O, pa, lg):
    """Natural order above 2^20 runs in three passes: two transforms at a distance beyond the
    footprint, pack 4, in place (one transform per chunk), forward and normalised inverse, against
    the oracle and against the four-step."""
    import torch

    n, batch, pack = 1 << lg, 2, 4
    dist = O.real_size(n, pack) + 64
    xs = [x64(O, 0x5EED0601 + b, n) for b in range(batch)]
    buf = np.zeros(dist * batch, np.float64)
    for b, x in enumerate(xs):
        pb = packed64(O, x, pack)
        buf[b * dist: b * dist + len(pb)] = pb
    outs = []
    for three in (True, False):
        d = dev(buf)
        unit = pa.fft_unit(n, dtype="float64")
        unit.set_three_pass(three)
        unit.set_chunk(16 * n)
        unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist)
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy())
        if three:
            h = outs[-1]
            for b, x in enumerate(xs):
                assert O.rel_l2(unpacked64(O, h[b * dist:], n, pack), O.fft64(x)) < tol64(n), b
                assert np.all(h[b * dist + O.real_size(n, pack): (b + 1) * dist] == 0)  # gaps untouched
            unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist, inverse=True)
            torch.cuda.synchronize()
            h = d.cpu().numpy()
            for b, x in enumerate(xs):
                assert O.rel_l2(unpacked64(O, h[b * dist:], n, pack), x) < 2 * tol64(n), b
    assert O.rel_l2(outs[0], outs[1]) < tol64(n)


def test_round_trip_f64_large(O, pa):
    """ifft(fft(x)) == x at 2^20 x 8 (a size-independent property; 128 MiB)."""
    import torch

    n, batch, pack = 1 << 20, 8, 4
    unit = pa.fft_unit(n, dtype="float64")
    dist = O.real_size(n, pack)
    d = torch.empty(dist * batch, dtype=torch.float64, device="cuda")
    from packed_amd import _lib

    # the batch is one packed vector of batch * n values (dist is exactly one footprint)
    _lib.check(_lib.lib().pcx_generate_f64(d.data_ptr(), pack, batch * n, 0x5EED0601, 1, 1.0, None))
    x0 = d.clone()
    unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist)
    unit.batch(d.data_ptr(), d.data_ptr(), batch, pack, dist, pack, dist, inverse=True)
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(d - x0) / torch.linalg.vector_norm(x0))
    assert err < 2 * tol64(n)


@pytest.mark.parametrize("lg", [4, 13, 16, 21])
def test_single_tone_closed_form_f64(O, pa, lg):
    """exp(2 pi i k j / N) has the spectrum N * delta_k (closed form, no oracle)."""
    import torch

    n, k = 1 << lg, (1 << lg) // 3 + 1
    j = np.arange(n)
    x = np.exp(2j * np.pi * ((k * j) % n) / n)
    unit = pa.fft_unit(n, dtype="float64")
    v = pa.vector.from_complex(x, dtype="float64")
    unit(v)
    torch.cuda.synchronize()
    want = np.zeros(n, np.complex128)
    want[k] = n
    assert O.rel_l2(v.to_complex(), want) < tol64(n)


def test_f64_type_checks(pa):
    from packed_amd import PcxError, _lib

    with pytest.raises(PcxError) as e:
        pa.fft_unit(1024, order="unordered", dtype="float64")
    assert e.value.code == _lib.ERR_UNSUPPORTED
    u32 = pa.fft_unit(1024)
    u64 = pa.fft_unit(1024, dtype="float64")
    v64 = pa.vector(1024, dtype="float64")
    v32 = pa.vector(1024)
    assert v64.pack == 4 and v32.pack == 8  # types.hpp:70-72
    with pytest.raises(PcxError):
        u32(v64)
    with pytest.raises(PcxError):
        u64(v32)
    # the raw C entry points refuse a plan of the other element type
    rc = _lib.lib().pcx_fft(u64.handle, v64.data_ptr(), 4, v64.data_ptr(), 4, 1024, None)
    assert rc == _lib.ERR_INVALID_ARGUMENT
    rc = _lib.lib().pcx_fft_f64(u32.handle, v64.data_ptr(), 4, v64.data_ptr(), 4, 1024, None)
    assert rc == _lib.ERR_INVALID_ARGUMENT
    assert _lib.lib().pcx_plan_element_size(u64.handle) == 8
    assert _lib.lib().pcx_plan_element_size(u32.handle) == 4


def test_f64_in_place_repack_message(pa):
    """fft.hpp:783, 889: a pack change wider than one SIMD register cannot run in place; a double
    register holds 4 values, so packs 4 -> 8 are refused for double (and allowed for float)."""
    from packed_amd import PcxError, _lib

    L = _lib.lib()
    u64 = pa.fft_unit(1024, dtype="float64")
    v64 = pa.vector(1024, pack=8, dtype="float64")
    rc = L.pcx_fft_f64(u64.handle, v64.data_ptr(), 8, v64.data_ptr(), 4, 1024, None)
    assert rc == _lib.ERR_INPLACE_REPACK
    assert L.pcx_last_error().decode() == ("cannot perform in-place repack from source pack size (which is 4) "
                                            "to destination pack size (which is 8)")
    rc = L.pcx_ifft_f64(u64.handle, 1, v64.data_ptr(), 2, v64.data_ptr(), 8, None)
    assert rc == _lib.ERR_INPLACE_REPACK
    u32 = pa.fft_unit(1024)
    v32 = pa.vector(1024, pack=8)
    assert L.pcx_fft(u32.handle, v32.data_ptr(), 8, v32.data_ptr(), 4, 1024, None) == _lib.SUCCESS
    with pytest.raises(PcxError):
        _lib.check(L.pcx_fft(u32.handle, v32.data_ptr(), 16, v32.data_ptr(), 8, 1024, None))
