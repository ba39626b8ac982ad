"""GPU parity of device-side vector expressions and pack conversion (pcx_vexpr_eval, pcx_repack).

Mirrors the reference's tests/test_op.cpp test_bin_ops (89-226): lhs[i] = (i eps, -1 - i),
rhs[i] = (i + 1, i eps); the 16 operations {+, -, *, /} x {a, conj(a)} x {b, conj(b)} between two
vectors and with a complex scalar (512(1 + eps), -512 eps) on either side, {+, -, *, /} with the
real scalar 1024(eps - 1) on either side; lengths 1 .. 31 and pack sizes 1 .. 32; every element
within the reference's equal_eps tolerance (10 FLT_EPSILON relative, tests/test_op.cpp:57-67) of
the FP64 result.  Pack conversion is checked bit for bit.
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _equal_eps(got, want):
    big = np.maximum(np.abs(got), np.abs(want))
    return np.all((got == want) | (np.abs(got - want) < big * 10 * EPS))


OPS = [("+", lambda a, b: a + b), ("-", lambda a, b: a - b), ("*", lambda a, b: a * b), ("/", lambda a, b: a / b)]


def _vec_ops(pa):
    c = pa.conj
    for (name, f), ca, cb in itertools.product(OPS, (False, True), (False, True)):
        def dev(a, b, f=f, ca=ca, cb=cb):
            return f(c(a) if ca else a, c(b) if cb else b)

        def host(a, b, f=f, ca=ca, cb=cb):
            return f(np.conj(a) if ca else a, np.conj(b) if cb else b)

        yield f"{'conj(a)' if ca else 'a'} {name} {'conj(b)' if cb else 'b'}", dev, host


@pytest.mark.parametrize("pack", [1, 2, 4, 8, 16, 32])
def test_reference_bin_ops(orc, pa, pack):
    import torch

    for length in list(range(1, 32)) + [1000]:
        i = np.arange(length, dtype=np.float64)
        lhs = (i * EPS + 1j * (-1.0 - i)).astype(np.complex64)
        rhs = ((i + 1) + 1j * (i * EPS)).astype(np.complex64)
        L = pa.vector.from_complex(lhs, pack)
        R = pa.vector.from_complex(rhs, pack)
        out = pa.vector(length, pack)
        lh, rh = lhs.astype(np.complex128), rhs.astype(np.complex128)
        cx = np.complex64(complex(512 * (EPS + 1), -512 * EPS))
        re = np.float32(1024 * (EPS - 1))
        for name, dev, host in _vec_ops(pa):
            pa.store(dev(L, R), out)
            torch.cuda.synchronize()
            assert _equal_eps(out.to_complex(), host(lh, rh)), (pack, length, name)
            # the same operation with a complex scalar on the right, then on the left
            # (conj of a scalar is taken on the host)
            pa.store(dev(L, complex(cx)), out)
            torch.cuda.synchronize()
            assert _equal_eps(out.to_complex(), host(lh, complex(cx))), (pack, length, name, "scalar rhs")
            pa.store(dev(complex(cx), R), out)
            torch.cuda.synchronize()
            assert _equal_eps(out.to_complex(), host(complex(cx), rh)), (pack, length, name, "scalar lhs")
        for (name, f) in OPS:
            for left in (False, True):
                for s, sh in ((complex(cx), complex(cx)), (float(re), float(re))):
                    e = f(s, L) if left else f(L, s)
                    pa.store(e, out)
                    torch.cuda.synchronize()
                    want = f(sh, lh) if left else f(lh, sh)
                    assert _equal_eps(out.to_complex(), want), (pack, length, name, left, s)


@pytest.mark.parametrize("packs", [(1, 8), (8, 1), (8, 8), (2, 32), (32, 4), (1, "N"), ("N", 8), ("N", 1)])
@pytest.mark.parametrize("n", [1, 17, 1000, 4096])
def test_repack_bit_exact(orc, pa, packs, n):
    import torch

    full = 1 << max(0, (n - 1).bit_length())  # "N": one pack holds the whole vector
    sp, dp = (full if p == "N" else p for p in packs)
    z = orc.gen_normal(0x5EED0050, n).astype(np.complex64)
    src = pa.vector.from_complex(z, sp)
    dst = pa.vector(n, dp)
    pa.repack(dst, src)
    torch.cuda.synchronize()
    got = dst.data.cpu().numpy()
    want = orc.to_packed(z, dp)
    assert np.array_equal(got[: len(want)], want)


def test_repack_in_place_rules(orc, pa):
    import torch

    z = orc.gen_normal(1, 100).astype(np.complex64)
    v = pa.vector.from_complex(z, 8)
    pa.repack(v, v)  # same pack: allowed, a no-op
    torch.cuda.synchronize()
    assert np.array_equal(v.to_complex(), z.astype(np.complex128))
    w = pa.vector(100, 1, data=v.data)  # the same memory seen with another pack
    with pytest.raises(pa.PcxError) as e:
        pa.repack(w, v)
    assert "in-place repack" in str(e.value)


def test_fused_expression_and_errors(orc, pa):
    """A deep expression in one pass (four registers) against FP64, and the error paths."""
    import torch

    n = 5000
    zs = [orc.gen_normal(0x5EED0060 + k, n).astype(np.complex64) for k in range(4)]
    packs = [1, 8, 64, 2]
    a, b, c, d = (pa.vector.from_complex(z, p) for z, p in zip(zs, packs))
    ah, bh, ch, dh = (z.astype(np.complex128) for z in zs)
    out = pa.vector(n, 16)
    e = ((a + b) * pa.conj(c - d)) / ((a - b) * (c + 3.0) + 4.0) - 2.5j * d
    pa.store(e, out)
    torch.cuda.synchronize()
    want = ((ah + bh) * np.conj(ch - dh)) / ((ah - bh) * (ch + 3.0) + 4.0) - 2.5j * dh
    assert orc.rel_l2(out.to_complex(), want) < 1e-6
    with pytest.raises(pa.PcxError):
        pa.store(a + pa.vector(n - 1, 8), out)  # size mismatch
    too_deep = ((a + b) * (c + d)) / ((a - b) * (c - d))
    too_deep = too_deep + ((a + b) * (c + d)) / ((a - b) * (c - d))
    with pytest.raises(pa.PcxError):
        pa.store(too_deep, out)


EPS64 = float(np.finfo(np.float64).eps)


@pytest.mark.parametrize("pack", [1, 4, 8, 32])
def test_reference_bin_ops_double(orc, pa, pack):
    """test_bin_ops for double (tests/test_op.cpp runs it for float and double): the same operands
    with DBL_EPSILON, every element within 10 DBL_EPSILON relative of an extended-precision
    (80-bit long double) host result."""
    import torch

    def eq(got, want):
        want = np.asarray(want, np.clongdouble)
        got = np.asarray(got, np.clongdouble)
        big = np.maximum(np.abs(got), np.abs(want))
        return np.all((got == want) | (np.abs(got - want) < big * 10 * EPS64))

    for length in list(range(1, 32)) + [1000]:
        i = np.arange(length, dtype=np.float64)
        lhs = i * EPS64 + 1j * (-1.0 - i)
        rhs = (i + 1) + 1j * (i * EPS64)
        L = pa.vector.from_complex(lhs, pack, dtype="float64")
        R = pa.vector.from_complex(rhs, pack, dtype="float64")
        out = pa.vector(length, pack, dtype="float64")
        lh, rh = lhs.astype(np.clongdouble), rhs.astype(np.clongdouble)
        cx = complex(512 * (EPS64 + 1), -512 * EPS64)
        re = 1024 * (EPS64 - 1)
        for name, dev, host in _vec_ops(pa):
            pa.store(dev(L, R), out)
            torch.cuda.synchronize()
            assert eq(out.to_complex(), host(lh, rh)), (pack, length, name)
            pa.store(dev(L, cx), out)
            torch.cuda.synchronize()
            assert eq(out.to_complex(), host(lh, np.clongdouble(cx))), (pack, length, name, "scalar rhs")
            pa.store(dev(cx, R), out)
            torch.cuda.synchronize()
            assert eq(out.to_complex(), host(np.clongdouble(cx), rh)), (pack, length, name, "scalar lhs")
        for (name, f) in OPS:
            for left in (False, True):
                for s in (cx, re):
                    pa.store(f(s, L) if left else f(L, s), out)
                    torch.cuda.synchronize()
                    sh = np.clongdouble(s) if isinstance(s, complex) else np.longdouble(s)
                    want = f(sh, lh) if left else f(lh, sh)
                    assert eq(out.to_complex(), want), (pack, length, name, left, s)


@pytest.mark.parametrize("sp,dp", [(1, 4), (4, 1), (4, 32), (2, 8)])
def test_repack_double_bit_exact(orc, pa, sp, dp):
    import torch

    n = 1000
    z = orc.gen_normal(0x5EED0051, n).astype(np.complex128) * (1 + 2.0 ** -40)
    src = pa.vector.from_complex(z, sp, dtype="float64")
    dst = pa.vector(n, dp, dtype="float64")
    pa.repack(dst, src)
    torch.cuda.synchronize()
    assert np.array_equal(dst.to_complex(), z)
    with pytest.raises(pa.PcxError):
        pa.repack(dst, pa.vector(n, sp))  # float and double do not mix
