"""Device-side vector expressions over packed complex vectors (vector_arithm.hpp:450-1130).

Mirrors the reference's expression operators: vector (+ - * /) vector for any two pack sizes,
vector (+ - * /) scalar and scalar (+ - * /) vector for a real (float) or complex scalar, and
conj().  An expression is lazy; `store(expr, dst)` (the reference's pcxo::store,
vector_arithm.hpp:405-447) flattens it into a program for pcx_vexpr_eval, which evaluates the
whole tree in one fused kernel on the GPU.

Register allocation: Sethi-Ullman numbering.  A node needs max(need(l), need(r)) registers when
the two differ and need + 1 when they are equal; the operand that needs more is evaluated first.
The machine has four registers, so any expression whose tree needs at most four evaluates.
"""
from __future__ import annotations

import numbers

from . import _lib
from ._lib import PcxError, check, lib

_BIN = {"+": "ADD", "-": "SUB", "*": "MUL", "/": "DIV"}
# scalar ops: (op with scalar on the right, op with scalar on the left); real and complex forms
_SCA = {"+": ("ADD", "ADD"), "-": ("SUB", "RSUB"), "*": ("MUL", "MUL"), "/": ("DIV", "RDIV")}


class Expr:
    """A lazy element-wise expression over packed device vectors."""

    def __add__(self, o):
        return _combine("+", self, o)

    def __radd__(self, o):
        return _combine("+", o, self)

    def __sub__(self, o):
        return _combine("-", self, o)

    def __rsub__(self, o):
        return _combine("-", o, self)

    def __mul__(self, o):
        return _combine("*", self, o)

    def __rmul__(self, o):
        return _combine("*", o, self)

    def __truediv__(self, o):
        return _combine("/", self, o)

    def __rtruediv__(self, o):
        return _combine("/", o, self)

    def size(self) -> int:
        raise NotImplementedError

    def need(self) -> int:
        raise NotImplementedError


class Leaf(Expr):
    def __init__(self, vec):
        self.vec = vec

    def size(self):
        return self.vec.size()

    def need(self):
        return 1


class Bin(Expr):
    def __init__(self, op, lhs, rhs):
        if lhs.size() != rhs.size():
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"vector sizes {lhs.size()} and {rhs.size()} differ")
        self.op, self.lhs, self.rhs = op, lhs, rhs

    def size(self):
        return self.lhs.size()

    def need(self):
        a, b = self.lhs.need(), self.rhs.need()
        return max(a, b) if a != b else a + 1


class Scalar(Expr):
    def __init__(self, op, expr, scalar, scalar_left):
        self.op, self.expr, self.scalar, self.left = op, expr, scalar, scalar_left

    def size(self):
        return self.expr.size()

    def need(self):
        return self.expr.need()


class Conj(Expr):
    def __init__(self, expr):
        self.expr = expr

    def size(self):
        return self.expr.size()

    def need(self):
        return self.expr.need()


def as_expr(x) -> Expr:
    if isinstance(x, Expr):
        return x
    if hasattr(x, "pack") and hasattr(x, "data_ptr"):
        return Leaf(x)
    raise TypeError(f"not a vector expression: {type(x).__name__}")


def _combine(op, a, b):
    a_s = isinstance(a, numbers.Number)
    b_s = isinstance(b, numbers.Number)
    if a_s and b_s:
        raise TypeError("an expression needs at least one vector")
    if b_s:
        return Scalar(op, as_expr(a), b, False)
    if a_s:
        return Scalar(op, as_expr(b), a, True)
    return Bin(op, as_expr(a), as_expr(b))


def conj(x):
    """vector_arithm.hpp:495 conj(expression); the conjugate of a scalar is the scalar's."""
    if isinstance(x, numbers.Number):
        return complex(x).conjugate() if isinstance(x, complex) else x
    return Conj(as_expr(x))


def compile_expr(expr: Expr):
    """Program (list of VInstr fields) and source list for pcx_vexpr_eval."""
    expr = as_expr(expr)
    if expr.need() > _lib.VEXPR_REGS:
        raise PcxError(_lib.ERR_UNSUPPORTED, f"expression needs {expr.need()} registers (at most 4)")
    srcs, index, prog = [], {}, []

    def src_of(vec):
        key = (vec.data_ptr(), vec.pack)
        if key not in index:
            index[key] = len(srcs)
            srcs.append(vec)
        return index[key]

    def emit(op, d, a=0, b=0, s=0j):
        prog.append((_lib.VOP[op], d, a, b, float(complex(s).real), float(complex(s).imag)))

    def gen(e, r):
        if isinstance(e, Leaf):
            emit("LOAD", r, src_of(e.vec))
        elif isinstance(e, Conj):
            gen(e.expr, r)
            emit("CONJ", r, r)
        elif isinstance(e, Scalar):
            gen(e.expr, r)
            right, left = _SCA[e.op]
            name = (left if e.left else right) + ("R" if isinstance(e.scalar, numbers.Real) else "S")
            emit(name, r, r, s=e.scalar)
        else:
            if e.lhs.need() >= e.rhs.need():
                gen(e.lhs, r)
                gen(e.rhs, r + 1)
                emit(_BIN[e.op], r, r, r + 1)
            else:
                gen(e.rhs, r)
                gen(e.lhs, r + 1)
                emit(_BIN[e.op], r, r + 1, r)

    gen(expr, 0)
    if len(prog) > _lib.VEXPR_MAX_PROG or len(srcs) > _lib.VEXPR_MAX_SRC:
        raise PcxError(_lib.ERR_UNSUPPORTED, "expression exceeds 64 instructions or 16 sources")
    return prog, srcs


def _common_dtype(*vs) -> str:
    """The element type shared by every vector of an evaluation (float32 or float64)."""
    kinds = {getattr(v, "dtype", "float32") for v in vs}
    if len(kinds) != 1:
        raise PcxError(_lib.ERR_INVALID_ARGUMENT, "an expression mixes float and double vectors")
    return kinds.pop()


def store(expr, dst, stream=None) -> None:
    """pcxo::store(expr, dst) (vector_arithm.hpp:405-447): evaluate into dst in its pack size."""
    import ctypes as C

    from .fft import _stream_ptr

    expr = as_expr(expr)
    if expr.size() != dst.size():
        raise PcxError(_lib.ERR_SIZE_MISMATCH,
                       f"expression size (which is {expr.size()}) differs from destination size (which is {dst.size()})")
    prog, srcs = compile_expr(expr)
    f64 = _common_dtype(dst, *srcs) == "float64"
    VS, VI = (_lib.VSource64, _lib.VInstr64) if f64 else (_lib.VSource, _lib.VInstr)
    S = (VS * max(1, len(srcs)))(*[VS(v.data_ptr(), v.pack) for v in srcs])
    P = (VI * len(prog))(*[VI(*i) for i in prog])
    ev = lib().pcx_vexpr_eval_f64 if f64 else lib().pcx_vexpr_eval
    check(ev(S, len(srcs), P, len(prog), dst.data_ptr(), dst.pack, dst.size(), _stream_ptr(stream)))


def repack(dst, src, stream=None) -> None:
    """Pack conversion dst <- src (same size, any two pack sizes)."""
    from .fft import _stream_ptr

    if src.size() != dst.size():
        raise PcxError(_lib.ERR_SIZE_MISMATCH, "repack needs equal sizes")
    rp = lib().pcx_repack_f64 if _common_dtype(dst, src) == "float64" else lib().pcx_repack
    check(rp(dst.data_ptr(), dst.pack, src.data_ptr(), src.pack, dst.size(), _stream_ptr(stream)))
