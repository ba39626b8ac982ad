"""Pins the CPU oracle (oracle/pcx_oracle.c) against checks that do not restate it:
numpy's pocketfft, closed forms, the C library's rand(), the reference's lane-shuffle network,
exact-integer correlations, and invariants chosen so a plausible mistake (wrong sign, dropped
twiddle, transposed index, off-by-one block) fails at least one of them.  The mutation tests at
the bottom plant such mistakes in copies of the oracle's numbers and show the checks catch them.
"""
import ctypes
import math

import numpy as np
import pytest

from helpers import fft_tol

SIZES = [2, 4, 8, 16, 32, 64, 128, 1024, 4096]


def test_fp64_fft_matches_numpy(orc):
    for n in SIZES + [1 << 16]:
        x = orc.gen_normal(n, n).astype(np.complex128)
        assert orc.rel_l2(orc.fft64(x), np.fft.fft(x)) < 1e-14
        assert orc.rel_l2(orc.fft64(x, True), np.fft.ifft(x) * n) < 1e-14


def test_brute_force_dft_matches_fft(orc):
    for n in SIZES:
        x = orc.gen_uniform(n + 1, n).astype(np.complex128)
        assert orc.rel_l2(orc.dft(x), orc.fft64(x)) < 1e-13
        assert orc.rel_l2(orc.dft(x, True), orc.fft64(x, True)) < 1e-13


def test_dft_closed_forms(orc):
    for n in [8, 64, 1024]:
        d = np.zeros(n, np.complex128)
        d[0] = 1
        assert np.allclose(orc.fft64(d), np.ones(n), atol=1e-15)  # delta -> all ones
        k0 = 5
        tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
        X = orc.fft64(tone)
        want = np.zeros(n, np.complex128)
        want[k0] = n  # a single tone lands in one bin; sign convention exp(-2*pi*i*nk/N)
        assert np.max(np.abs(X - want)) < 1e-10 * n
        # leaky tone of tests/test_fft.cpp:299-302, closed-form geometric sum
        f = 13.37
        x = np.exp(1j * 2 * np.pi * np.arange(n) / n * f)
        dk = 2j * np.pi * (f - np.arange(n))
        assert orc.rel_l2(orc.fft64(x), (1 - np.exp(dk)) / (1 - np.exp(dk / n))) < 1e-12


def test_dft_invariants(orc):
    n = 512
    a = orc.gen_normal(1, n).astype(np.complex128)
    b = orc.gen_normal(2, n).astype(np.complex128)
    A, B = orc.fft64(a), orc.fft64(b)
    assert orc.rel_l2(orc.fft64(2.5 * a - 1j * b), 2.5 * A - 1j * B) < 1e-14  # linearity
    assert abs(np.sum(np.abs(A) ** 2) / n - np.sum(np.abs(a) ** 2)) < 1e-9 * n  # Parseval
    r = np.real(a)
    R = orc.fft64(r)
    assert np.allclose(R[1:], np.conj(R[1:][::-1]), atol=1e-10)  # real input -> Hermitian
    assert orc.rel_l2(orc.fft64(orc.fft64(a), True) / n, a) < 1e-14  # round trip


def test_glibc_rand_restatement(orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(1)
    g = orc.GlibcRand(1)
    assert [g() for _ in range(3000)] == [libc.rand() for _ in range(3000)]


def test_unordered_map_two_derivations(orc):
    """Closed form q vs the replay of avx2_fft.hpp:188-213's register shuffles."""
    for i in range(256):
        assert orc.unordered_q(i) == orc.unordered_q_network(i)
    assert sorted(orc.unordered_q(i) for i in range(32)) == list(range(32))  # a permutation
    assert orc.unordered_q(1) == 4 and orc.unordered_q(8) == 1  # b0 -> bit 2, b3 -> bit 0


def test_order_maps(orc):
    for n in [32, 64, 1024]:
        bits = int(math.log2(n))
        br = orc.slot_freq("bit_reversed", n)
        assert all(br[i] == int(format(i, f"0{bits}b")[::-1], 2) for i in range(n))
        un = orc.slot_freq("unordered", n)
        assert sorted(un) == list(range(n)) and not np.array_equal(un, br)
        assert np.array_equal(orc.slot_freq("normal", n), np.arange(n))
        for order in ["normal", "bit_reversed", "unordered"]:  # array form == scalar form
            assert list(orc.slot_freq(order, n)) == [orc.lib().orc_slot_freq(orc.ORDERS[order], i, n)
                                                     for i in range(n)]


def test_pack_layout(orc):
    """element_access.hpp:8-11 / vector.hpp:373-381."""
    assert [orc.lib().orc_pidx(i, 8) for i in range(10)] == [0, 1, 2, 3, 4, 5, 6, 7, 16, 17]
    assert orc.lib().orc_real_size(9, 8) == 32 and orc.lib().orc_real_size(16, 1) == 32
    z = orc.gen_uniform(3, 37)
    for p in [1, 2, 8, 64]:
        assert np.array_equal(orc.from_packed(orc.to_packed(z, p), 37, p), z.astype(np.complex128))


@pytest.mark.parametrize("order", ["normal", "bit_reversed", "unordered"])
def test_fft_unit_semantics(orc, order):
    n = 256
    x = orc.gen_uniform(9, n)
    X = np.fft.fft(x.astype(np.complex128))
    f = orc.slot_freq(order, n)
    for sp, dp in [(8, 8), (1, 8), (8, 1), (256, 4)]:
        dst, vals = orc.fft_unit(order, False, True, n, orc.to_packed(x, sp), sp, dp)
        assert orc.rel_l2(vals, X[f]) < 1e-14
        spk = orc.spectrum_pack(order, dp)
        assert orc.rel_l2(orc.from_packed(dst, n, spk), X[f]) < 1e-7
        back, xv = orc.fft_unit(order, True, True, n, dst, dp, sp)
        assert orc.rel_l2(xv, x) < 1e-7
        _, xu = orc.fft_unit(order, True, False, n, dst, dp, sp)
        assert orc.rel_l2(xu, n * x) < 1e-7
    # zero-extension: the source is read up to src_size only
    _, vz = orc.fft_unit(order, False, True, n, orc.to_packed(x, 8), 8, 8, 100)
    xz = x.astype(np.complex128).copy()
    xz[100:] = 0
    assert orc.rel_l2(vz, np.fft.fft(xz)[f]) < 1e-14


def test_golden_vectors(orc):
    """tests/golden/fft_n1024.npz: config-1 input (N=1024, uniform, seed 0x5EED0001) and its spectra in
    the three orders, pack 8, computed by numpy's pocketfft (FP64) and the closed-form index maps."""
    import os

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_n1024.npz"))
    n = int(g["n"])
    x = orc.gen_uniform(int(g["seed"]), n)
    assert np.array_equal(orc.to_packed(x, 8), g["input_pack8"])
    for order in ["normal", "bit_reversed", "unordered"]:
        dst, vals = orc.fft_unit(order, False, True, n, g["input_pack8"], 8, 8)
        want = g[f"{order}_pack8"]
        assert np.all(np.abs(dst - want) <= np.spacing(np.abs(want))), order  # within 1 ulp of FP32
        assert orc.rel_l2(vals, g[f"{order}_c128"]) < 1e-14


def test_xcorr_barker_exact(orc):
    """Barker-13 aperiodic autocorrelation: 13 at lag 0, sidelobes 0 or 1 (test_fxcorr.cpp:34-48 data)."""
    b = np.array([1, 1, 1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1], dtype=np.float64)
    f = np.zeros(128)
    f[35:48] = b
    out = orc.xcorr_naive(f, b)
    assert out[35] == 13
    side = np.delete(np.real(out[23:48]), 12)
    assert set(np.abs(side).tolist()) <= {0.0, 1.0}


def test_fxcorr_overlap_save_restatement(orc):
    for L, M, N, pack in [(5000, 100, 256, 8), (4096, 64, 512, 8), (777, 33, 64, 1), (100, 1, 8, 8)]:
        f = orc.gen_normal(L, L).astype(np.complex128)
        g = orc.gen_normal(M, M).astype(np.complex128)
        a = orc.xcorr_naive(f, g)
        b, nb = orc.fxcorr_overlap_save(f, g, N, pack)
        step = (N - (M - 1)) // pack * pack
        assert nb == -(-L // step)
        assert orc.rel_l2(b, a) < 1e-12


def test_port_fp32(orc):
    """The FP32 host-core port: within the FP32 bound of FP64, every order, both directions."""
    for n in [32, 64, 512, 4096, 1 << 15]:
        p = orc.PortPlan(n)
        x = orc.gen_uniform(n, n)
        X = np.fft.fft(x.astype(np.complex128))
        for order in ["normal", "bit_reversed", "unordered"]:
            y = p.fft(x, order)
            assert orc.rel_l2(y, X[orc.slot_freq(order, n)]) < fft_tol(n)
            assert orc.rel_l2(p.fft(y, order, inverse=True), x) < 2 * fft_tol(n)
    f = orc.gen_normal(5, 20000).astype(np.complex64)
    g = orc.gen_normal(6, 500).astype(np.complex64)
    got, nb = orc.port_fxcorr(f, g, 2048, threads=4)
    assert orc.rel_l2(got, orc.xcorr_naive(f, g)) < 4e-6


# ---------------------------------------------------------------------- mutation checks
def test_checks_catch_planted_mistakes(orc):
    """Each plausible oracle mistake, planted in a copy of its output, fails a check above."""
    n = 64
    x = orc.gen_uniform(11, n).astype(np.complex128)
    X = orc.fft64(x)
    ref = np.fft.fft(x)
    assert orc.rel_l2(X, ref) < 1e-14
    wrong_sign = np.fft.ifft(x) * n                      # exp(+2*pi*i) instead of exp(-2*pi*i)
    transposed = ref[orc.slot_freq("bit_reversed", n)]   # order map applied twice / forgotten
    dropped_twiddle = np.concatenate([np.fft.fft(x[0::2]), np.fft.fft(x[0::2])]) + \
        np.concatenate([np.fft.fft(x[1::2]), -np.fft.fft(x[1::2])])  # last radix-2 stage without W
    off_by_one = np.roll(ref, 1)
    for bad in [wrong_sign, transposed, dropped_twiddle, off_by_one]:
        assert orc.rel_l2(bad, ref) > 1e-3
    # a q map with two bits swapped differs from the shuffle-network replay
    bad_q = [(i & 16) | ((i & 4) << 1) | ((i & 2) << 1) | ((i & 1) << 1) | ((i >> 3) & 1) for i in range(32)]
    assert bad_q != [orc.unordered_q_network(i) for i in range(32)]
    # a block step off by one leaves an output wrong in the multi-block correlation
    f = orc.gen_normal(12, 3000).astype(np.complex128)
    g = orc.gen_normal(13, 50).astype(np.complex128)
    good, _ = orc.fxcorr_overlap_save(f, g, 128, 8)
    bad, _ = orc.fxcorr_overlap_save(f, g, 128, 1)  # different step, same answer (step rounding is free)
    assert orc.rel_l2(bad, good) < 1e-12
    a = orc.xcorr_naive(f, g)
    shifted = np.concatenate([a[1:], [0]])
    assert orc.rel_l2(shifted, a) > 1e-3
