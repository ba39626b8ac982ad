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


# ---- the permuted exchange of row_group / k_dif (bit_reversed and unordered loads and stores)
def plan(la, f64):
    a = 1 << la
    if f64:
        ppt = 32 if la >= 13 else (16 if la >= 4 else a)
    else:
        ppt = 64 if la >= 14 else 32 if la >= 13 else (16 if la >= 4 else a)
    return ppt, a // ppt


def brev(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2) if bits else 0


def uq(i):  # the unordered lane permutation of the low five bits
    return (i & ~31) | (i & 16) | ((i & 4) << 1) | ((i & 1) << 2) | (i & 2) | ((i >> 3) & 1)


def perm_base(jt, la, ppt, ta, unordered):
    lt, lpp = ta.bit_length() - 1, ppt.bit_length() - 1
    return brev(uq(jt) if unordered else jt, lt) << lpp


CASES = [(la, False) for la in range(1, 15)] + [(la, True) for la in range(1, 14)]


@pytest.mark.parametrize("la,f64", CASES)
def test_permuted_padding_is_carry_free_and_conflict_free(la, f64):
    a = 1 << la
    ppt, ta = plan(la, f64)
    pp = max(la - 6, 5)
    pad_p = lambda x: x + (x >> pp)  # noqa: E731
    stride = a + (a >> 5) + 1  # the row stride the LDS array is sized for
    assert pad_p(a - 1) < stride
    rs = radices(la, False)
    for rad in (rs[0], rs[-1]):  # first-stage loads, last-stage stores: b = jt + m*TA, o = r*A/R
        b = (np.arange(ta)[:, None] + np.arange(ppt // rad)[None, :] * ta).reshape(-1, 1)
        o = np.arange(rad)[None, :] * (a // rad)
        assert np.array_equal(pad_p(b + o), pad_p(b) + o + (o >> pp))
    if ta >= 32:  # permuted side: b = perm_base(jt), o = brev(u) < PPT
        lpp = ppt.bit_length() - 1
        for unordered in (False, True):
            if unordered and la < 5:
                continue
            b = np.array([perm_base(j, la, ppt, ta, unordered) for j in range(ta)])[:, None]
            o = np.array([brev(u, lpp) for u in range(ppt)])[None, :]
            assert np.array_equal(pad_p(b + o), pad_p(b) + o + (o >> pp))
            if ta >= 64:  # one wave = 64 lanes of one transform: distinct banks per access
                # (from 2^11 on; at 2^10 the pad stays one per 32 elements, which the row stride
                # is sized for, and two lanes can share a bank)
                need = 64 if la >= 11 else 32
                for w0 in range(0, ta, 64):
                    banks = pad_p(b[w0:w0 + 64] + o) % 64
                    assert all(len(set(banks[:, u])) >= need for u in range(ppt))


def test_old_padding_conflicts_in_permuted_order():
    """With one pad per 32 elements the permuted reads of a 2^14-point FP32 plan hit 8 banks:
    the check above would catch a return to that layout."""
    la, (ppt, ta) = 14, plan(14, False)
    b = np.array([perm_base(j, la, ppt, ta, False) for j in range(64)])
    assert len(set((b + (b >> 5)) % 64)) == 8


# ---- the 2^14 correlation layout with butterflies in permuted thread order (fxcorr_kernels.hpp)
def xl_lds(e):
    return ((e & 0x300F) | ((e >> 4) & 0xF0) | ((e << 4) & 0xF00)) + ((e >> 4) & 0xF) + ((e >> 8) & 0x30)


def xl_jt(t):
    return (t & ~0x330) | ((t >> 4) & 0x30) | ((t << 4) & 0x300)


def xl_accesses():
    """(name, base[thread], offsets) of the LDS exchanges of the 2^14 plan (radices 16, 16, 16, 4)
    that stay in LDS: stage-0 writes, stage-1 reads, stage-1 writes, stage-2 reads."""
    jt = xl_jt(np.arange(1024))
    r = np.arange(16)
    return [
        ("write0", jt * 16, r),                          # Ns = 1: hi = 16 j, offset r
        ("read1", jt, r * 1024),                         # j + r * A/16
        ("write1", (jt // 16) * 256 + (jt & 15), r * 16),  # Ns = 16
        ("read2", jt, r * 1024),
    ]


def test_xl_layout_is_injective_and_fits_the_row():
    e = np.arange(1 << 14)
    a = xl_lds(e)
    assert len(np.unique(a)) == a.size
    assert a.max() < (1 << 14) + (1 << 9) + 1  # the row stride the LDS array is sized for


def test_xl_layout_is_additive_and_conflict_free():
    for name, base, offs in xl_accesses():
        full = base[:, None] + offs[None, :]
        assert np.array_equal(xl_lds(full), xl_lds(base)[:, None] + xl_lds(offs)[None, :]), name
        for w0 in range(0, 1024, 64):  # 64 lanes of one wavefront on 64 distinct banks
            banks = xl_lds(full[w0:w0 + 64]) % 64
            assert all(len(set(banks[:, u])) == 64 for u in range(16)), (name, w0)


def test_xl_order_needs_its_layout():
    """The plain padded layout with the same thread order puts up to four lanes on one bank."""
    worst = 64
    for name, base, offs in xl_accesses():
        full = base[:, None] + offs[None, :]
        banks = pad(full[:64]) % 64
        worst = min(worst, min(len(set(banks[:, u])) for u in range(16)))
    assert worst <= 16
