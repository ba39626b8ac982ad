"""Writes tests/golden/fft_n1024.npz: the N=1024 pack-8 workload (re/im uniform [-1, 1) from
splitmix64 seed 0x5EED0001) and its forward spectra in the three fft_order values.

The expected outputs come from numpy's pocketfft in FP64 (not from the oracle) and the closed
form index maps: normal slot i = X[i], bit_reversed slot i = X[brev(i)], unordered slot i =
X[brev(q(i))] with q(b4 b3 b2 b1 b0) = (b4 b2 b0 b1 b3) on the low five bits.  The reference
library cannot be compiled in this environment (fft.hpp:15 includes <bits/utility.h>, which the
installed libstdc++ 11 does not ship), so its outputs are not part of these vectors.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402  (input generator and packing only)

N = 1024
SEED = 0x5EED0001


def brev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2)


def q(i):
    b = [(i >> k) & 1 for k in range(5)]
    return (i & ~31) | (b[4] << 4) | (b[2] << 3) | (b[0] << 2) | (b[1] << 1) | b[3]


def main():
    x = O.gen_uniform(SEED, N)
    X = np.fft.fft(x.astype(np.complex128))
    bits = 10
    maps = {
        "normal": np.arange(N),
        "bit_reversed": np.array([brev(i, bits) for i in range(N)]),
        "unordered": np.array([brev(q(i), bits) for i in range(N)]),
    }
    out = {"n": np.int64(N), "seed": np.int64(SEED), "input_pack8": O.to_packed(x, 8)}
    for name, m in maps.items():
        out[f"{name}_c128"] = X[m]
        out[f"{name}_pack8"] = O.to_packed(X[m].astype(np.complex64), 8)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_n1024.npz"), **out)


if __name__ == "__main__":
    main()
