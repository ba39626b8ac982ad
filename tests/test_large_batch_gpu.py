"""GPU parity at the largest batch extents: more than 2^32 work-items in one call, and batches
whose element offsets pass 2^31 (the reference's fft_unit batch loops, fft.hpp:565-915, and the
overlap-save loop over blocks, fxcorr.hpp:107-147, have no such limit).

A dispatch covers at most 2^32 - 1 work-items, so the library caps the grid of its grid-stride
kernels and splits the correlation into launches of whole blocks.  These cases check every
transform (or output) of such calls through size-independent properties on the device, and
sampled transforms (outputs) against the oracle's FP64 definition:

* 2^32 + 24 two-point transforms in one call (64 GiB): the closed form X0 = x0 + x1,
  X1 = x0 - x1 over the whole batch;
* 2^16-point four-step batches of 8 GiB + 3 transforms (FP32 and FP64): several scratch chunks
  and a partial one, transforms past element offset 2^31; forward checked per sampled
  transform, then the in-place inverse against the input over the whole batch;
* a correlation of 2^30 + 2^20 samples with step 1 (64-point blocks, 64 taps): 2^32 + 2^22
  work-items in two launches; output windows at both ends and around 2^30 against the direct
  sum out[i] = sum_m f[i+m] conj(g[m]);
* a correlation of 2^28 + 12345 samples (pack 8) with 16384-point blocks and 4096 taps, whose
  footprint passes 2^29 floats: the 2^14 block kernel with 64-bit element offsets, windows at
  both ends, around 2^28 and at block seams against the same direct sum.

Device memory peaks at 128 GiB (the first case).
"""
import gc

import numpy as np
import pytest

from helpers import EPS32, fft_tol

pytestmark = pytest.mark.gpu

EPS64 = 2.0 ** -52


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _free():
    import torch

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _gen(dst, pack, n, seed, f64=False):
    from packed_amd import _lib

    fn = _lib.lib().pcx_generate_f64 if f64 else _lib.lib().pcx_generate
    _lib.check(fn(dst.data_ptr(), pack, n, seed, 0, 1.0, None))


def test_two_point_batch_beyond_2pow32_work_items(orc, pa):
    """2^32 + 24 transforms of 2 points, pack 2, one thread each: the single-pass grid exceeds
    2^32 work-items and is capped; the kernel's grid-stride loop covers the rest."""
    import torch

    batch = (1 << 32) + 24
    x = torch.empty(4 * batch, dtype=torch.float32, device="cuda")  # [re0 re1 im0 im1] per transform
    y = torch.empty_like(x)
    _gen(x, 2, 2 * batch, 0x5EED0701)  # one packed vector of 2*batch values has the batch layout
    unit = pa.fft_unit(2)
    unit.batch(y.data_ptr(), x.data_ptr(), batch, 2, 4, 2, 4)
    torch.cuda.synchronize()
    xv, yv = x.view(batch, 4), y.view(batch, 4)
    worst, rows = 0.0, 1 << 27
    for r0 in range(0, batch, rows):
        a, b = xv[r0:r0 + rows], yv[r0:r0 + rows]
        want = torch.stack([a[:, 0] + a[:, 1], a[:, 0] - a[:, 1], a[:, 2] + a[:, 3], a[:, 2] - a[:, 3]], 1)
        worst = max(worst, float((b - want).abs().max()))
        del a, b, want
    assert worst <= 2.0 * EPS32, worst  # |x| <= 1: one rounding of a sum below 2
    for t in (0, (1 << 31) - 1, 1 << 31, (1 << 32) - 1, 1 << 32, batch - 1):
        xs, ys = xv[t].cpu().numpy().astype(np.float64), yv[t].cpu().numpy().astype(np.float64)
        want = orc.fft64(np.array([xs[0] + 1j * xs[2], xs[1] + 1j * xs[3]]))
        assert np.allclose([ys[0] + 1j * ys[2], ys[1] + 1j * ys[3]], want, rtol=0, atol=4 * EPS32), t
    del x, y, xv, yv, unit
    _free()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_four_step_batch_beyond_8_gib(orc, pa, dtype):
    """2^16-point transforms, 8 GiB + 3 transforms out of place: scratch chunks of 2 GiB, a last
    chunk of 3 transforms, and transforms whose elements lie past offset 2^31."""
    import torch

    f64 = dtype == "float64"
    n, pack = 1 << 16, (4 if f64 else 8)
    full = (8 << 30) // (2 * n * (8 if f64 else 4))  # 8 GiB of transforms
    batch = full + 3
    dist = 2 * n
    tdt = torch.float64 if f64 else torch.float32
    tol = 4 * 16 * EPS64 if f64 else fft_tol(n)
    x = torch.empty(dist * batch, dtype=tdt, device="cuda")
    y = torch.empty_like(x)
    _gen(x, pack, n * batch, 0x5EED0702, f64)
    unit = pa.fft_unit(n, dtype=dtype)
    unit.batch(y.data_ptr(), x.data_ptr(), batch, pack, dist, pack, dist)
    torch.cuda.synchronize()
    chunk = (2048 << 20) // (n * (16 if f64 else 8))  # transforms per scratch chunk
    for b in (0, chunk - 1, chunk, full - 1, full, batch - 1):
        xs = x[b * dist:(b + 1) * dist].cpu().numpy().astype(np.float64)
        ys = y[b * dist:(b + 1) * dist].cpu().numpy().astype(np.float64)
        got = orc.from_packed(ys, n, pack)
        want = orc.fft64(orc.from_packed(xs, n, pack))
        assert orc.rel_l2(got, want) < tol, (b, orc.rel_l2(got, want))
    # inverse in place over the whole batch: back to the input
    unit.batch(y.data_ptr(), y.data_ptr(), batch, pack, dist, pack, dist, inverse=True)
    torch.cuda.synchronize()
    num = den = 0.0
    step = dist * (1 << 12)
    for o in range(0, dist * batch, step):
        a, b = x[o:o + step].double(), y[o:o + step].double()
        num += float(((b - a) ** 2).sum())
        den += float((a ** 2).sum())
        del a, b
    assert (num / den) ** 0.5 < 2 * tol
    del x, y, unit
    _free()


def test_correlation_beyond_2pow32_work_items(orc, pa):
    """Step-1 overlap-save (N = 64, M = 64, pack 1): one 64-point block per output sample,
    4 threads each, 2^30 + 2^20 blocks in launches of under 2^32 work-items."""
    import torch

    n_fft, m = 64, 64
    length = (1 << 30) + (1 << 20)
    g = orc.gen_normal(0x5EED0703, m)
    gv = pa.vector.from_complex(g, 1)
    xc = pa.fxcorr_unit(gv, n_fft)
    f = torch.empty(2 * length, dtype=torch.float32, device="cuda")
    out = torch.empty_like(f)
    _gen(f, 1, length, 0x5EED0704)
    xc.batch(out.data_ptr(), f.data_ptr(), length, 1, 1, 2 * length)
    torch.cuda.synchronize()
    w = 4096
    rng = np.random.default_rng(0x5EED0705)
    starts = [0, (1 << 30) - w, 1 << 30, length - w] + [int(s) for s in rng.integers(0, length - w, 4)]
    for a in starts:
        hi = min(a + w + m, length)
        fs = f[2 * a:2 * hi].cpu().numpy().astype(np.float64)
        fw = fs[0::2] + 1j * fs[1::2]
        want = orc.xcorr_naive(fw, g)[:w]  # exact for the first w outputs (f is zero past L)
        os_ = out[2 * a:2 * (a + w)].cpu().numpy().astype(np.float64)
        got = os_[0::2] + 1j * os_[1::2]
        assert orc.rel_l2(got, want) < 64 * EPS32, (a, orc.rel_l2(got, want))
    del f, out, xc, gv
    _free()


def test_correlation_2p14_blocks_beyond_32bit_offsets(orc, pa):
    """N = 16384, M = 4096, pack 8, L = 2^28 + 12345: the signal's packed footprint passes 2^29
    floats, so the blocks take 64-bit element offsets (fxcorr_kernels.hpp, B32 = false)."""
    import torch

    n_fft, m, pack = 16384, 4096, 8
    length = (1 << 28) + 12345
    g = orc.gen_normal(0x5EED0706, m)
    gv = pa.vector.from_complex(g, pack)
    xc = pa.fxcorr_unit(gv, n_fft)
    step = xc.step(pack)
    foot = 2 * (-(-length // pack)) * pack
    f = torch.empty(foot, dtype=torch.float32, device="cuda")
    out = torch.empty_like(f)
    _gen(f, pack, length, 0x5EED0707)
    xc.batch(out.data_ptr(), f.data_ptr(), length, 1, pack, foot)
    torch.cuda.synchronize()
    w = 2048
    seams = [k * step - w // 2 for k in (1, 7, (1 << 28) // step)]
    starts = [0, (1 << 28) - w, length - w] + seams
    for a in starts:
        a -= a % pack  # whole packs, so the host slices below stay aligned
        hi = min(a + w + m, length)
        fw = orc.from_packed(f[2 * a:2 * (a + (-(-(hi - a) // pack)) * pack)].cpu().numpy(), hi - a, pack)
        want = orc.xcorr_naive(fw.astype(np.complex128), g)[:w]  # exact: f is zero past L
        got = orc.from_packed(out[2 * a:2 * (a + w)].cpu().numpy(), w, pack)
        assert orc.rel_l2(got, want) < 64 * EPS32, (a, orc.rel_l2(got, want))
    del f, out, xc, gv
    _free()
