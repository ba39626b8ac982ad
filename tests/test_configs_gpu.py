"""GPU parity of the benchmark workloads at their stated sizes.

* Correlation of 2^24 samples with 16 copies of a 4096-tap unit-modulus kernel at SNR 0 dB, FFT
  block 16384: the 16 peaks index-exact, and output windows at both ends of the signal and
  around block seams against the direct FP64 definition (tests/test_fxcorr.cpp:14-32).
* 512 channels of 2^20 samples correlated with one 1024-tap kernel, in FP32 and both split
  precisions, FFT blocks 4096 and 8192: sampled channels (first, middle, last) against the
  direct definition at the start, at block seams and at the end; the whole batch in place equals
  the whole batch out of place.
* 10,000 transforms of log-uniform sizes 2^6 .. 2^20 with mixed directions in one work queue: the
  first and last transform of every (size, direction) class against the FP64 oracle, and the
  whole batch through a second queue of the opposite directions back to its input.
The correlation's reference is only correct for one block (fxcorr.hpp:112-129 offsets a float
pointer), so multi-block outputs are checked against the definition it encodes.
"""
import math

import numpy as np
import pytest

from helpers import fft_tol

pytestmark = pytest.mark.gpu

XC_TOL = 4e-6  # relative L2 of FP32 overlap-save against the FP64 direct sum


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _window_err(orc, f, got, g, a, b):
    """Relative L2 of got[a:b] against the direct correlation of f (zero beyond its end)."""
    M = len(g)
    want = orc.xcorr_naive(f[a:min(len(f), b + M - 1)], g)[: b - a]
    return orc.rel_l2(got[a:b], want)


def test_fxcorr_2p24_peaks_and_windows(orc, pa):
    import torch

    L, M, N = 1 << 24, 4096, 16384
    f = orc.gen_normal(0x5EED0003, L, math.sqrt(0.5))
    ph = orc.gen_uniform(0x5EED0013, M)
    g = np.exp(1j * np.pi * (np.real(ph).astype(np.float64) + 1.0)).astype(np.complex64)
    rng = np.random.default_rng(0x5EED0003)
    offs = np.sort(rng.choice(np.arange(0, L - M, 2 * M), 16, replace=False))
    for o in offs:
        f[o:o + M] += g
    xc = pa.fxcorr_unit(pa.vector.from_complex(g, 8), N)
    src = pa.vector.from_complex(f, 8)
    out = pa.vector(L, 8)
    xc(src, out)
    torch.cuda.synchronize()
    got = out.to_complex()
    top = np.sort(np.argsort(np.abs(got))[-16:])
    assert np.array_equal(top, offs), (top, offs)
    step = xc.step(8)
    nblocks = -(-L // step)
    f64 = f.astype(np.complex128)
    g64 = g.astype(np.complex128)
    wins = [(0, 4096), (L - 4096, L)]
    for k in (1, 2, nblocks // 2, nblocks - 1):
        wins.append((k * step - 512, min(L, k * step + 512)))
    for a, b in wins:
        assert _window_err(orc, f64, got, g64, a, b) < XC_TOL, (a, b)
    # in place, bit for bit the same outputs
    xc(src)
    torch.cuda.synchronize()
    assert torch.equal(src.data, out.data)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x2"])
@pytest.mark.parametrize("N", [4096, 8192])
def test_fxcorr_512_channels(orc, pa, precision, N):
    import torch

    from packed_amd._lib import check, lib

    C, L, M, pack = 512, 1 << 20, 1024, 8
    dist = 2 * L
    st = torch.cuda.current_stream().cuda_stream
    sig = torch.empty(C * dist, dtype=torch.float32, device="cuda")
    check(lib().pcx_generate(sig.data_ptr(), pack, C * L, 0x5EED0015, 1, math.sqrt(0.5), st))
    out = torch.empty_like(sig)
    g = np.exp(1j * np.pi * (np.real(orc.gen_uniform(0x5EED0016, M)).astype(np.float64) + 1.0)).astype(np.complex64)
    xc = pa.fxcorr_unit(pa.vector.from_complex(g, pack), N, precision=precision)
    xc.batch(out.data_ptr(), sig.data_ptr(), L, C, pack, dist)
    torch.cuda.synchronize()
    step = xc.step(pack)
    nblocks = -(-L // step)
    g64 = g.astype(np.complex128)
    for c in (0, C // 2, C - 1):
        f = orc.from_packed(sig[c * dist:(c + 1) * dist].cpu().numpy(), L, pack).astype(np.complex128)
        got = orc.from_packed(out[c * dist:(c + 1) * dist].cpu().numpy(), L, pack)
        wins = [(0, 2048), (L - 2048, L)]
        for k in (1, 2, nblocks // 2, nblocks - 1):
            wins.append((k * step - 256, min(L, k * step + 256)))
        for a, b in wins:
            assert _window_err(orc, f, got, g64, a, b) < XC_TOL, (precision, N, c, a, b)
    xc.batch(sig.data_ptr(), sig.data_ptr(), L, C, pack, dist)  # in place
    torch.cuda.synchronize()
    assert torch.equal(sig, out)


def _uneven_jobs(seed=0x5EED0004, count=10000):
    """The uneven-batch job list: sizes 2^6 .. 2^20 log-uniform, mixed directions, the vectors laid
    out grouped by (size, direction) in split layout (re[N] then im[N])."""
    rng = np.random.default_rng(seed)
    lgs = rng.integers(6, 21, count)
    invs = rng.integers(0, 2, count)
    order = np.lexsort((invs, -lgs))
    jobs, off = [], 0
    for k in order:
        n = 1 << int(lgs[k])
        jobs.append((int(lgs[k]), int(invs[k]), off, off))
        off += 2 * n
    return jobs, off


def test_uneven_10000_jobs(orc, pa):
    import torch

    from packed_amd._lib import check, lib

    jobs, total = _uneven_jobs()
    assert len(jobs) == 10000
    st = torch.cuda.current_stream().cuda_stream
    src = torch.empty(total, dtype=torch.float32, device="cuda")
    check(lib().pcx_generate(src.data_ptr(), 1, total // 2, 0x5EED0004, 1, math.sqrt(0.5), st))
    dst = torch.empty_like(src)
    q = pa.fft_queue(jobs)
    q.exec(dst.data_ptr(), src.data_ptr())
    torch.cuda.synchronize()
    classes = {}
    for lg, inv, so, do in jobs:
        classes.setdefault((lg, inv), []).append((so, do))
    for (lg, inv), members in sorted(classes.items()):
        n = 1 << lg
        for so, do in (members[0], members[-1]):
            xin = src[so:so + 2 * n].cpu().numpy()
            z = xin[:n].astype(np.float64) + 1j * xin[n:]
            want = orc.fft64(z, inverse=bool(inv)) / (n if inv else 1)
            xo = dst[do:do + 2 * n].cpu().numpy()
            err = orc.rel_l2(xo[:n] + 1j * xo[n:].astype(np.float64), want)
            assert err < fft_tol(n), (lg, inv, err)
    # the whole batch back through the opposite directions
    back = pa.fft_queue([(lg, 1 - inv, so, do) for lg, inv, so, do in jobs])
    rt = torch.empty_like(src)
    back.exec(rt.data_ptr(), dst.data_ptr())
    torch.cuda.synchronize()
    rt -= src
    err = float(torch.linalg.vector_norm(rt) / torch.linalg.vector_norm(src))
    assert err < 2 * fft_tol(1 << 20), err
