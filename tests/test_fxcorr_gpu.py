"""GPU parity of the overlap-save cross-correlation (fxcorr.hpp:44-147).

The definition is the one the reference test encodes (tests/test_fxcorr.cpp:14-32):
out[i] = sum_{m<M} f[i+m] * conj(g[m]), f zero beyond L.  The reference itself is only correct
for a single block (its block offset is applied to a float pointer, fxcorr.hpp:112-129), so
multi-block results are checked against the FP64 definition, not against the reference.
"""
import numpy as np
import pytest

from helpers import EPS32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


BARKER13 = np.array([1, 1, 1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1], dtype=np.float64)


def _xc_tol(f, g):
    # FP32 overlap-save: error relative to |f|*|g| energy per output, generous multiple of eps*log2(N)
    return 64 * EPS32


def _run(pa, f, g, fft_size, pack=8, inplace=True):
    import torch

    gv = pa.vector.from_complex(g, pack)
    xc = pa.fxcorr_unit(gv, fft_size)
    fv = pa.vector.from_complex(f, pack)
    if inplace:
        xc(fv)
        out = fv
    else:
        out = pa.vector(len(f), pack)
        xc(fv, out)
    torch.cuda.synchronize()
    return out.to_complex(), xc


def test_reference_barker_case(orc, pa):
    """tests/test_fxcorr.cpp:100-139: g = Barker-13, f = i*(Barker at 0 and 109) + (rand()%7 - 3),
    N = 256.  Every value is an integer combination, so the expected output is exact."""
    size = 128
    g = np.zeros(size)
    g[:13] = BARKER13
    sf = np.zeros(size, np.complex128)
    sf[:13] += BARKER13
    sf[109:122] += BARKER13
    rnd = orc.GlibcRand(1)
    for i in range(size):
        sf[i] = sf[i] * 1j + (rnd() % 7 - 3)
    want = orc.xcorr_naive(sf, g)
    assert np.array_equal(np.round(want), want)  # exact integers
    assert want[0].imag == 13 and want[109].imag == 13  # Barker peaks
    got, _ = _run(pa, sf, g, size * 2)
    assert np.max(np.abs(got - want)) < 1e-4 * np.max(np.abs(want))


def test_reference_random_phase_case(orc, pa):
    """tests/test_fxcorr.cpp:61-98 test_xcorr<float>(512): M = 512 random phases exp(i*(rand()%6-3)),
    N = 1024, L = 512, with rand() continuing after the Barker case's 128 draws."""
    rnd = orc.GlibcRand(1)
    for _ in range(128):
        rnd()
    g = np.array([np.exp(1j * (rnd() % 6 - 3)) for _ in range(512)])
    f = np.array([np.exp(1j * (rnd() % 6 - 3)) for _ in range(512)])
    want = orc.xcorr_naive(f, g)
    got, _ = _run(pa, f.astype(np.complex64), g.astype(np.complex64), 1024)
    # reference tolerance is k = 1e7 FLT_EPSILON per element (tests/test_fxcorr.cpp:91); ours is tighter
    assert orc.rel_l2(got, want) < 2e-6


@pytest.mark.parametrize("L,M,N", [(5000, 100, 256), (4096, 64, 512), (70000, 4096, 16384), (12345, 1, 64),
                                   (3000, 57, 128), (1 << 16, 1024, 4096), (5000, 200, 256), (40000, 9000, 16384)])
@pytest.mark.parametrize("inplace", [True, False])
def test_multi_block(orc, pa, L, M, N, inplace):
    f = orc.gen_normal(0x5EED0003, L, np.sqrt(0.5)).astype(np.complex128)
    ph = orc.gen_uniform(0x5EED0013, M)
    g = np.exp(1j * np.pi * np.real(ph))
    want = orc.xcorr_naive(f, g)
    got, xc = _run(pa, f.astype(np.complex64), g.astype(np.complex64), N, inplace=inplace)
    assert xc.step(8) == (N - (M - 1)) // 8 * 8
    assert orc.rel_l2(got, want) < 4e-6, orc.rel_l2(got, want)


def test_overlap_save_oracle_agrees(orc):
    """The FP64 overlap-save restatement equals the direct definition (pins the block logic)."""
    f = orc.gen_normal(7, 9000).astype(np.complex128)
    g = orc.gen_normal(8, 300).astype(np.complex128)
    a = orc.xcorr_naive(f, g)
    b, nb = orc.fxcorr_overlap_save(f, g, 1024, 8)
    assert nb == -(-9000 // ((1024 - 299) // 8 * 8))
    assert orc.rel_l2(b, a) < 1e-13


def test_peaks_index_exact(orc, pa):
    """Scaled-down form of the 2^24-sample workload: 16 copies of a 4096-tap unit-modulus kernel
    buried at SNR 0 dB; the 16 largest outputs sit exactly at the insertion offsets."""
    L, M, N = 1 << 20, 4096, 16384
    rng_ph = orc.gen_uniform(0x5EED0004, M)
    g = np.exp(1j * np.pi * np.real(rng_ph))
    f = orc.gen_normal(0x5EED0003, L, np.sqrt(0.5)).astype(np.complex128)
    offs = np.sort(np.random.default_rng(0x5EED0003).choice(np.arange(0, L - M, M), 16, replace=False))
    for o in offs:
        f[o:o + M] += g
    got, _ = _run(pa, f.astype(np.complex64), g.astype(np.complex64), N)
    top = np.sort(np.argsort(np.abs(got))[-16:])
    assert np.array_equal(top, offs)
    want = orc.xcorr_naive(f[: 1 << 16], g)[: (1 << 16) - M]
    assert orc.rel_l2(got[: (1 << 16) - M], want) < 4e-6


@pytest.mark.parametrize("inplace", [True, False])
def test_multi_channel(orc, pa, inplace):
    """pcx_fxcorr_exec_batch: several channels correlated with one kernel spectrum in one launch."""
    import torch

    L, M, N, C = 20000, 300, 1024, 5
    pack = 8
    dist = orc.real_size(L, pack) + 64
    g = np.exp(1j * np.pi * np.real(orc.gen_uniform(0x5EED0013, M)))
    fs = [orc.gen_normal(0x5EED0030 + c, L, np.sqrt(0.5)).astype(np.complex128) for c in range(C)]
    buf = np.zeros(dist * C, np.float32)
    for c in range(C):
        pb = orc.to_packed(fs[c].astype(np.complex64), pack)
        buf[c * dist:c * dist + len(pb)] = pb
    src = torch.from_numpy(buf).cuda()
    dst = src if inplace else torch.zeros_like(src)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g.astype(np.complex64), pack), N)
    xc.batch(dst.data_ptr(), src.data_ptr(), L, C, pack, dist)
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    for c in range(C):
        got = orc.from_packed(out[c * dist:], L, pack)
        assert orc.rel_l2(got, orc.xcorr_naive(fs[c], g)) < 4e-6, c


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x2"])
def test_split_correlation(orc, pa, precision):
    """Split-precision matrix-core forms of the fused correlation."""
    import torch

    L, M, N = 70000, 4096, 16384
    f = orc.gen_normal(0x5EED0003, L, np.sqrt(0.5)).astype(np.complex128)
    g = np.exp(1j * np.pi * np.real(orc.gen_uniform(0x5EED0013, M)))
    gv = pa.vector.from_complex(g.astype(np.complex64), 8)
    xc = pa.fxcorr_unit(gv, N, precision=precision)
    fv = pa.vector.from_complex(f.astype(np.complex64), 8)
    out = pa.vector(L, 8)
    xc(fv, out)
    torch.cuda.synchronize()
    assert orc.rel_l2(out.to_complex(), orc.xcorr_naive(f, g)) < 4e-6


def test_multichannel_distance_is_checked(orc, pa):
    """Channels closer than one signal's footprint (or at a negative distance) are refused before
    any launch."""
    import torch

    from packed_amd import PcxError

    g = pa.vector.from_complex(orc.gen_uniform(0x5EED0F00, 64))
    xc = pa.fxcorr_unit(g, 512)
    buf = torch.zeros(4 * 2 * 1000, dtype=torch.float32, device="cuda")
    for dist in (1000, -2000):  # footprint of 1000 values at pack 8 is 2000 floats
        with pytest.raises(PcxError):
            xc.batch(buf.data_ptr(), buf.data_ptr(), 1000, 2, 8, dist)
    xc.batch(buf.data_ptr(), buf.data_ptr(), 1000, 2, 8, 2000)  # exactly one footprint apart: fine
    torch.cuda.synchronize()


@pytest.mark.parametrize("L,M,N", [(200000, 20000, 1 << 15), (300000, 5000, 1 << 16), (70000, 30000, 1 << 17)])
@pytest.mark.parametrize("inplace", [True, False])
def test_block_path_above_16384(orc, pa, L, M, N, inplace):
    """FFT blocks above 16384 (fxcorr.hpp:62-69 takes any power of two) run through HBM and the
    batched four-step FFT; kernels longer than 16384 taps included."""
    f = orc.gen_normal(0x5EED0061, L, np.sqrt(0.5)).astype(np.complex128)
    g = np.exp(1j * np.pi * np.real(orc.gen_uniform(0x5EED0062, M)))
    got, xc = _run(pa, f.astype(np.complex64), g.astype(np.complex64), N, inplace=inplace)
    assert xc.step(8) == (N - (M - 1)) // 8 * 8
    assert orc.rel_l2(got, orc.xcorr_naive(f, g)) < 4e-6


@pytest.mark.parametrize("L,M,N,C", [(20000, 300, 1024, 3), (3000, 57, 128, 4), (150000, 20000, 1 << 15, 2),
                                     (100000, 10, 64, 2), (30000, 150, 256, 3)])
@pytest.mark.parametrize("inplace", [True, False])
def test_workspace_matches_own_buffers(orc, pa, L, M, N, C, inplace):
    """pcx_fxcorr_exec_batch_ws (the TmpFactory_ slot, fxcorr.hpp:10-32): with a workspace of one
    fft_size vector, or three, the outputs equal the unit's own-buffer path bit for bit."""
    import torch

    pack = 8
    dist = orc.real_size(L, pack)
    g = np.exp(1j * np.pi * np.real(orc.gen_uniform(0x5EED0063, M)))
    buf = np.zeros(dist * C, np.float32)
    for c in range(C):
        buf[c * dist:(c + 1) * dist] = orc.to_packed(orc.gen_normal(0x5EED0064 + c, L, np.sqrt(0.5)), pack)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g.astype(np.complex64), pack), N)
    assert xc.workspace_min() == 8 * N
    src = torch.from_numpy(buf).cuda()
    ref = src.clone() if inplace else torch.zeros_like(src)
    xc.batch(ref.data_ptr(), (ref if inplace else src).data_ptr(), L, C, pack, dist)
    for mult in (1, 3):
        ws = torch.empty(2 * N * mult, dtype=torch.float32, device="cuda")
        got = src.clone() if inplace else torch.zeros_like(src)
        xc.batch(got.data_ptr(), (got if inplace else src).data_ptr(), L, C, pack, dist, workspace=ws.data_ptr(),
                 workspace_bytes=ws.numel() * 4)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), mult
    fs = [orc.from_packed(buf[c * dist:], L, pack).astype(np.complex128) for c in range(C)]
    out = ref.cpu().numpy()
    for c in range(C):
        assert orc.rel_l2(orc.from_packed(out[c * dist:], L, pack), orc.xcorr_naive(fs[c], g)) < 4e-6


def test_workspace_too_small_is_refused(orc, pa):
    import torch

    from packed_amd import PcxError

    g = pa.vector.from_complex(orc.gen_uniform(0x5EED0F01, 64))
    xc = pa.fxcorr_unit(g, 512)
    buf = torch.zeros(2 * 1000, dtype=torch.float32, device="cuda")
    ws = torch.empty(2 * 512, dtype=torch.float32, device="cuda")
    with pytest.raises(PcxError):
        xc.batch(buf.data_ptr(), buf.data_ptr(), 1000, 1, 8, 2000, workspace=ws.data_ptr(), workspace_bytes=4 * 1023)
    pa.fxcorr_unit(g, 1024, precision="bf16x3")  # split precision covers fft_size 16 .. 16384
    with pytest.raises(PcxError):
        pa.fxcorr_unit(g, 1 << 15, precision="bf16x3")


@pytest.mark.parametrize("pack", [1, 64, 1024])
@pytest.mark.parametrize("inplace", [True, False])
def test_2p14_blocks_every_pack(orc, pa, pack, inplace):
    """16384-point blocks (the kernel whose exchange before the radix-4 stage runs in registers)
    at packs 1, 64 and 1024 (the step, floor((N - M + 1) / P) * P, stays a whole number of packs:
    fxcorr.hpp:62-69), three channels, in place (saved overlaps, per-element loads) and out of
    place."""
    import torch

    L, M, N, C = 90000, 5000, 16384, 3
    dist = orc.real_size(L, pack) + 2 * pack
    g = orc.gen_normal(0x5EED0940, M)
    fs = [orc.gen_normal(0x5EED0941 + c, L, np.sqrt(0.5)).astype(np.complex128) for c in range(C)]
    buf = np.zeros(dist * C, np.float32)
    for c in range(C):
        pb = orc.to_packed(fs[c].astype(np.complex64), pack)
        buf[c * dist:c * dist + len(pb)] = pb
    src = torch.from_numpy(buf).cuda()
    dst = src if inplace else torch.zeros_like(src)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g, pack), N)
    xc.batch(dst.data_ptr(), src.data_ptr(), L, C, pack, dist)
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    for c in range(C):
        got = orc.from_packed(out[c * dist:], L, pack)
        assert orc.rel_l2(got, orc.xcorr_naive(fs[c], g.astype(np.complex128))) < 4e-6, (pack, c)


@pytest.mark.parametrize("N,M", [(16384, 4096), (8192, 1024), (256, 30)])
def test_in_place_block_order_repeats_and_replays(orc, pa, N, M):
    """In place with N - step <= step the blocks run in ticket order and the unit's order words
    are left zero by every launch: calls of growing and shrinking length on one unit, then a
    captured graph replayed three times, each equal bit for bit to the out-of-place result."""
    import torch

    pack = 8
    g = np.exp(1j * np.pi * np.real(orc.gen_uniform(0x5EED0071, M))).astype(np.complex64)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g, pack), N)
    step = xc.step(pack)
    assert N - step <= step
    for L in (40 * step + 17, 7 * step, 90 * step + 3, 3 * step + 1):
        f = orc.gen_normal(0x5EED0072 + L, L, np.sqrt(0.5)).astype(np.complex64)
        src = pa.vector.from_complex(f, pack)
        oop = pa.vector(L, pack)
        xc(src, oop)
        xc(src)
        torch.cuda.synchronize()
        assert torch.equal(src.data, oop.data), L
    L = 60 * step + 5
    f = orc.gen_normal(0x5EED0073, L, np.sqrt(0.5)).astype(np.complex64)
    x0 = pa.vector.from_complex(f, pack)
    want = pa.vector(L, pack)
    xc(x0, want)
    v = pa.vector.from_complex(f, pack)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        graph.capture_begin()
        xc(v, stream=s)
        graph.capture_end()
    torch.cuda.synchronize()
    for _ in range(3):
        v.data.copy_(x0.data)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(v.data, want.data)
