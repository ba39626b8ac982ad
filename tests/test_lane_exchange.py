"""CPU check of the in-register exchange of packed_amd/csrc/stockham.hpp (lane_exchange).

For plans with 16 points per thread, a small last radix RL and TA <= 64 threads per transform,
the Stockham exchange before the last stage (stage_write of the second-to-last stage, then
stage_read of the last one, through LDS) is replaced by swapping lane bits
[log2(TA/RL), log2 TA) with register bits [0, log2 RL).  This replays both on index arrays and
checks that every lane ends with the same value in every register, for each size the kernels
switch it on for, and for FP32 and FP64 plans alike (same index algebra).  It also models the
cross-lane primitives as lane permutations (v_permlane32_swap / v_permlane16_swap trade the odd
half-wave / rows of their first operand with the even ones of the second) and checks that each
realises the single-bit swap it is used for.
"""
import numpy as np
import pytest

ON = [9, 10]  # lane_exchange_on() in stockham.hpp


def radices(la):
    nfull, rem = la // 4, la % 4
    return [16] * nfull + ([1 << rem] if rem else [])


def lds_exchange(la, vals):
    """vals[lane, q] after the second-to-last stage -> values after stage_read of the last."""
    a, ppt = 1 << la, 16
    ta = a // ppt
    rs = radices(la)
    s = len(rs) - 2
    r, ns = rs[s], int(np.prod(rs[:s])) if s else 1
    lds = np.full(a, -1, dtype=np.int64)
    for jt in range(ta):  # stage_write of stage s
        for m in range(ppt // r):
            j = jt + m * ta
            hi = (j // ns) * ns * r + (j & (ns - 1))
            for k in range(r):
                lds[hi + k * ns] = vals[jt, m * r + k]
    assert (lds >= 0).all()
    rl = rs[-1]
    out = np.empty_like(vals)
    for jt in range(ta):  # stage_read of the last stage
        for m in range(ppt // rl):
            for k in range(rl):
                out[jt, m * rl + k] = lds[jt + m * ta + k * (a // rl)]
    return out


def lane_exchange(la, vals):
    a = 1 << la
    ta = a // 16
    lt, lr = ta.bit_length() - 1, radices(la)[-1].bit_length() - 1
    out = vals.copy()
    for i in range(lr):
        lb = lt - lr + i
        new = out.copy()
        for lane in range(ta):
            for q in range(16):
                # lane l, register q takes what lane l ^ 2^lb held in register q ^ 2^i, when the
                # lane bit and the register bit differ
                if (lane >> lb & 1) != (q >> i & 1):
                    new[lane, q] = out[lane ^ (1 << lb), q ^ (1 << i)]
        out = new
    return out


@pytest.mark.parametrize("la", ON + [5, 6, 7])
def test_register_exchange_matches_lds(la):
    """(2^5 .. 2^7 swap lane bits 0..2, which the kernels leave to LDS; the algebra is the same.)"""
    ta = (1 << la) // 16
    assert ta <= 64 and radices(la)[-1] < 16
    vals = np.arange(ta * 16).reshape(ta, 16)
    assert np.array_equal(lane_exchange(la, vals), lds_exchange(la, vals))


def test_a_wrong_bit_pairing_is_caught():
    """Pairing lane bit 4 of a 2^10 plan with register bit 1 (instead of 0) differs."""
    la = 10
    vals = np.arange(64 * 16).reshape(64, 16)
    want = lds_exchange(la, vals)
    bad = vals.copy()
    for lb, i in ((4, 1), (5, 0)):
        new = bad.copy()
        for lane in range(64):
            for q in range(16):
                if (lane >> lb & 1) != (q >> i & 1):
                    new[lane, q] = bad[lane ^ (1 << lb), q ^ (1 << i)]
        bad = new
    assert not np.array_equal(bad, want)


# ---- the cross-lane primitives as lane permutations of one 64-lane register
def permlane_swap(a, b, width):
    """v_permlane32_swap (width 32) / v_permlane16_swap (width 16): odd blocks of `a` trade
    places with even blocks of `b`; returns (new a, new b) = (builtin result [0], [1])."""
    a, b = a.copy(), b.copy()
    for base in range(0, 64, 2 * width):
        lo, hi = slice(base, base + width), slice(base + width, base + 2 * width)
        a[hi], b[lo] = b[lo].copy(), a[hi].copy()
    return a, b


@pytest.mark.parametrize("lb", [4, 5])
def test_primitives_swap_one_lane_bit(lb):
    a = np.arange(64) * 2
    b = np.arange(64) * 2 + 1  # value = 2 * lane + register bit
    na, nb = permlane_swap(a, b, 1 << lb)
    for lane in range(64):
        for q, reg in ((0, na), (1, nb)):
            # expected: the value of (lane with bit lb := q, register := lane's bit lb)
            src_lane = (lane & ~(1 << lb)) | (q << lb)
            assert reg[lane] == 2 * src_lane + (lane >> lb & 1), (lb, lane, q)


# ---- 2^14 correlation blocks (fxcorr_kernels.hpp, xc_lanes): TA = 1024 threads per transform
def xl_jt(t):
    """Butterfly index of thread t: t with bits 4, 5 and 8, 9 swapped (fxcorr_kernels.hpp)."""
    return (t & ~0x330) | ((t >> 4) & 0x30) | ((t << 4) & 0x300)


def test_xl_jt_places_index_bits_8_9_in_lane_bits_4_5():
    t = np.arange(1024)
    jt = xl_jt(t)
    assert sorted(jt) == list(range(1024))  # a bijection
    assert np.array_equal(jt >> 8 & 3, t >> 4 & 3)
    for lb, ib in ((4, 8), (5, 9)):  # flipping the lane bit flips the index bit, nothing else
        assert np.array_equal(xl_jt(t ^ (1 << lb)), jt ^ (1 << ib))


def permlane_exchange(vals, jt_of):
    """Swap lane bits 4, 5 with register bits 0, 1 by permlane swaps inside each wavefront, with
    thread t holding butterfly index jt_of(t); returns the result indexed by butterfly index."""
    ta = vals.shape[0]
    t = np.arange(ta)
    phys = vals[jt_of(t)]  # indexed by thread
    for i, lb in enumerate((4, 5)):
        new = phys.copy()
        for q in range(16):
            if q >> i & 1:
                continue
            a, b = phys[:, q].copy(), phys[:, q | 1 << i].copy()
            for base in range(0, ta, 64):
                w = slice(base, base + 64)
                new[w, q], new[w, q | 1 << i] = permlane_swap(a[w], b[w], 1 << lb)
        phys = new
    out = np.empty_like(phys)
    out[jt_of(t)] = phys
    return out


def test_register_exchange_2_14_on_physical_lanes():
    """The 2^14 plan (radices 16, 16, 16, 4): the LDS exchange before the radix-4 stage equals
    permlane swaps of lane bits 4 and 5 with register bits 0 and 1 when thread t holds butterfly
    index xl_jt(t)."""
    la, ta = 14, 1024
    vals = np.arange(ta * 16).reshape(ta, 16)
    want = lds_exchange(la, vals)
    assert np.array_equal(lane_exchange(la, vals), want)  # the algebra, in butterfly-index space
    assert np.array_equal(permlane_exchange(vals, xl_jt), want)
    # with thread t holding index t, index bits 8, 9 are wave bits: the same swaps are wrong
    assert not np.array_equal(permlane_exchange(vals, lambda t: t), want)
