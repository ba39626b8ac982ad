"""CPU checks of the vector-expression compiler (packed_amd/vexpr.py) and of the oracle's
restatement of the per-element machine (oracle.vexpr_run).

The oracle is pinned against numpy's own complex arithmetic on well-conditioned data (its
division formula is the reference's naive (a conj(b)) / |b|^2, simd_common.hpp:296-309), and the
compiler is pinned by running its programs through the oracle and comparing with a direct numpy
evaluation of the same expression trees, including trees that need all four registers."""
import numpy as np
import pytest

from packed_amd import vexpr as V
from packed_amd._lib import PcxError


class FakeVec:
    """Host stand-in with the attributes the compiler reads (no device memory involved)."""

    def __init__(self, z, pack):
        self.z = np.asarray(z, np.complex128)
        self.pack = pack
        self._p = id(self)

    def size(self):
        return len(self.z)

    def data_ptr(self):
        return self._p


def _direct(e):
    if isinstance(e, V.Leaf):
        return e.vec.z
    if isinstance(e, V.Conj):
        return np.conj(_direct(e.expr))
    if isinstance(e, V.Scalar):
        x, s = _direct(e.expr), e.scalar
        a, b = (s, x) if e.left else (x, s)
        return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[e.op]
    a, b = _direct(e.lhs), _direct(e.rhs)
    return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[e.op]


def _random_tree(rng, leaves, depth):
    if depth == 0 or rng.random() < 0.2:
        return V.Leaf(leaves[rng.integers(len(leaves))])
    k = rng.integers(4)
    if k == 0:
        return V.conj(_random_tree(rng, leaves, depth - 1))
    if k == 1:
        s = [2.5, -0.75, 1.5 - 0.5j, 0.25j][rng.integers(4)]
        op = "+-*/"[rng.integers(4)]
        e = _random_tree(rng, leaves, depth - 1)
        return V._combine(op, s, e) if rng.random() < 0.5 else V._combine(op, e, s)
    op = "+-*/"[rng.integers(4)]
    return V.Bin(op, _random_tree(rng, leaves, depth - 1), _random_tree(rng, leaves, depth - 1))


def test_oracle_division_matches_numpy(orc):
    rng = np.random.default_rng(1)
    a = rng.normal(size=1000) + 1j * rng.normal(size=1000)
    b = rng.normal(size=1000) + 1j * rng.normal(size=1000) + 3
    assert np.max(np.abs(orc.cdiv_ref(a, b) - a / b) / np.abs(a / b)) < 1e-14


def test_compiled_programs_match_direct_evaluation(orc):
    rng = np.random.default_rng(2)
    leaves = [FakeVec(rng.normal(size=64) + 1j * rng.normal(size=64) + 2.0, p) for p in (1, 8, 64, 2)]
    tested = 0
    for _ in range(300):
        e = _random_tree(rng, leaves, 5)
        if e.need() > 4:
            with pytest.raises(PcxError):
                V.compile_expr(e)
            continue
        prog, srcs = V.compile_expr(e)
        with np.errstate(all="ignore"):
            got = orc.vexpr_run(prog, [s.z for s in srcs])
            want = _direct(e)
        if not np.all(np.isfinite(want)) or np.max(np.abs(want)) > 1e6:
            continue  # a random tree that divides by (x - x) or overflows
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9)
        tested += 1
    assert tested > 150


def test_register_need_and_source_sharing():
    a, b, c, d = (FakeVec(np.ones(4), p) for p in (1, 2, 4, 8))
    e = (V.Leaf(a) + V.Leaf(b)) * (V.Leaf(c) + V.Leaf(d))
    assert e.need() == 3
    prog, srcs = V.compile_expr(e * V.Leaf(a))
    assert len(srcs) == 4  # a is loaded from one source slot twice
    deep = ((V.Leaf(a) + V.Leaf(b)) * (V.Leaf(c) + V.Leaf(d))) / ((V.Leaf(a) - V.Leaf(b)) * (V.Leaf(c) - V.Leaf(d)))
    assert deep.need() == 4
    V.compile_expr(deep)
    deeper = deep + ((V.Leaf(a) + V.Leaf(b)) * (V.Leaf(c) + V.Leaf(d))) / ((V.Leaf(a) - V.Leaf(b)) * (V.Leaf(c) - V.Leaf(d)))
    assert deeper.need() == 5
    with pytest.raises(PcxError):
        V.compile_expr(deeper)


def test_real_scalar_semantics(orc):
    """simd_common.hpp:390-410: a real scalar adds to / subtracts from the real part only, and the
    machine's real-scalar ops equal complex ops with a zero imaginary part on finite data."""
    x = np.array([1 + 2j, -3 - 0.5j])
    for name, f in [("ADDR", lambda v: v + 2), ("SUBR", lambda v: v - 2), ("RSUBR", lambda v: 2 - v),
                    ("MULR", lambda v: v * 2), ("DIVR", lambda v: v / 2), ("RDIVR", lambda v: 2 / v)]:
        from packed_amd._lib import VOP
        prog = [(0, 0, 0, 0, 0.0, 0.0), (VOP[name], 0, 0, 0, 2.0, 0.0)]
        assert np.allclose(orc.vexpr_run(prog, [x]), f(x), rtol=1e-15)
