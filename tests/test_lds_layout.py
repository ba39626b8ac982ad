"""CPU check of the padded LDS layout assumption of packed_amd/csrc/stockham.hpp.

Every Stockham stage addresses LDS as addr(base, off) = pad(base) + off + (off >> 5) with a
compile-time offset, which equals pad(base + off) only if adding `off` never carries across a
32-float padding boundary differently.  This enumerates every read and write of every plan
(LA = 1 .. 14, forward and reversed radix order) and checks the identity exactly.
"""
import numpy as np
import pytest


def pad(x):
    return x + (x >> 5)


def radices(la, rev):
    if la < 4:
        return [1 << la]
    nfull, rem = la // 4, la % 4
    rs = [16] * nfull
    if rem:
        rs = ([1 << rem] + rs) if rev else (rs + [1 << rem])
    return rs


@pytest.mark.parametrize("la", list(range(1, 15)))
@pytest.mark.parametrize("rev", [False, True])
def test_padding_is_carry_free(la, rev):
    a = 1 << la
    ns = 1
    for r in radices(la, rev):
        j = np.arange(a // r)[:, None]
        rr = np.arange(r)[None, :]
        o = rr * (a // r)  # stage reads: j + r * A/R
        assert np.array_equal(pad(j + o), pad(j) + o + (o >> 5))
        hi = (j // ns) * ns * r + (j & (ns - 1))  # stage writes: hi + r * Ns
        o = rr * ns
        assert np.array_equal(pad(hi + o), pad(hi) + o + (o >> 5))
        ns *= r


def test_a_carrying_layout_is_caught():
    """The same check rejects a layout that does carry (pad every 16 floats with offsets of 24)."""
    j = np.arange(64)[:, None]
    o = np.arange(4)[None, :] * 24
    p16 = lambda x: x + (x >> 4)  # noqa: E731
    assert not np.array_equal(p16(j + o), p16(j) + o + (o >> 4))
