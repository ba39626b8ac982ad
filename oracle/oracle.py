"""ctypes front end of the CPU oracle (liboracle.so, built from pcx_oracle.c).

Test infrastructure only: tests/, the smoke check and bench.py's cpu_baseline leg import this
module as the checker.  The product package never imports it.

Layout helpers follow the reference's packed layout: complex i of a pack-P vector has its real
part at float offset ``i + (i // P) * P`` and its imaginary part ``P`` floats later
(element_access.hpp:8-11, vector.hpp:373-381).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORDERS = {"normal": 0, "bit_reversed": 1, "unordered": 2}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i64, u64, f32p, f64p = C.c_int64, C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.orc_gen_uniform.argtypes = [u64, f32p, f32p, i64]
        L.orc_gen_normal.argtypes = [u64, C.c_double, f32p, f32p, i64]
        L.orc_brev.argtypes = [i64, C.c_int]
        L.orc_brev.restype = i64
        L.orc_unordered_q.argtypes = [i64]
        L.orc_unordered_q.restype = i64
        L.orc_unordered_q_network.argtypes = [i64]
        L.orc_unordered_q_network.restype = i64
        L.orc_slot_freq.argtypes = [C.c_int, i64, i64]
        L.orc_slot_freq.restype = i64
        L.orc_slot_map.argtypes = [C.c_int, i64, C.POINTER(i64)]
        L.orc_pidx.argtypes = [i64, i64]
        L.orc_pidx.restype = i64
        L.orc_real_size.argtypes = [i64, i64]
        L.orc_real_size.restype = i64
        L.orc_dft.argtypes = [f64p, f64p, f64p, f64p, i64, C.c_int]
        L.orc_fft.argtypes = [f64p, f64p, f64p, f64p, i64, C.c_int]
        L.orc_fft_unit.argtypes = [C.c_int, C.c_int, C.c_int, i64, f32p, i64, f32p, i64, i64, f64p, f64p]
        L.orc_xcorr_naive.argtypes = [f64p, f64p, i64, f64p, f64p, i64, f64p, f64p]
        L.orc_fxcorr_overlap_save.argtypes = [f64p, f64p, i64, f64p, f64p, i64, i64, i64, f64p, f64p]
        L.orc_fxcorr_overlap_save.restype = i64
        L.orc_glibc_srand.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_glibc_rand.argtypes = [C.c_void_p]
        L.orc_glibc_rand.restype = C.c_int32
        L.orc_port_plan_create.argtypes = [i64]
        L.orc_port_plan_create.restype = C.c_void_p
        L.orc_port_plan_destroy.argtypes = [C.c_void_p]
        L.orc_port_fft.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, f32p, f32p]
        L.orc_port_fft_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, f32p, f32p, i64, i64, C.c_int]
        L.orc_port_fxcorr.argtypes = [f32p, f32p, i64, f32p, f32p, i64, i64, f32p, f32p, C.c_int]
        L.orc_port_fxcorr.restype = i64
        _lib = L
    return _lib


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ inputs
def gen_uniform(seed: int, n: int) -> np.ndarray:
    """complex64 vector, re and im i.i.d. uniform [-1, 1) from splitmix64(seed)."""
    re = np.empty(n, np.float32)
    im = np.empty(n, np.float32)
    lib().orc_gen_uniform(seed, _f32(re), _f32(im), n)
    return (re + 1j * im).astype(np.complex64)


def gen_normal(seed: int, n: int, sigma: float = 1.0) -> np.ndarray:
    """complex64 vector, re and im i.i.d. normal(0, sigma^2) (Box-Muller over splitmix64)."""
    re = np.empty(n, np.float32)
    im = np.empty(n, np.float32)
    lib().orc_gen_normal(seed, sigma, _f32(re), _f32(im), n)
    return (re + 1j * im).astype(np.complex64)


class GlibcRand:
    """glibc rand() restated (TYPE_3 additive feedback), default seed 1 like an unseeded std::rand()."""

    def __init__(self, seed: int = 1):
        self._st = (C.c_int32 * 36)()
        lib().orc_glibc_srand(C.byref(self._st), seed)

    def __call__(self) -> int:
        return int(lib().orc_glibc_rand(C.byref(self._st)))


# ------------------------------------------------------------------ layout
def pidx(i, pack: int):
    i = np.asarray(i, dtype=np.int64)
    return i + (i // pack) * pack


def real_size(n: int, pack: int) -> int:
    return ((n + pack - 1) // pack) * pack * 2


def to_packed(z: np.ndarray, pack: int, n_alloc: int | None = None) -> np.ndarray:
    """complex vector -> float32 buffer in pack-P layout (zero padding up to whole packs)."""
    n = len(z) if n_alloc is None else n_alloc
    buf = np.zeros(real_size(n, pack), np.float32)
    idx = pidx(np.arange(len(z)), pack)
    buf[idx] = np.real(z)
    buf[idx + pack] = np.imag(z)
    return buf


def from_packed(buf: np.ndarray, n: int, pack: int) -> np.ndarray:
    idx = pidx(np.arange(n), pack)
    return (buf[idx].astype(np.float64) + 1j * buf[idx + pack].astype(np.float64))


def spectrum_pack(order: str, pack: int) -> int:
    """Pack of an unordered spectrum is max(P, 8) (avx2_fft.hpp:209-213, 227)."""
    return max(pack, 8) if order == "unordered" else pack


# ------------------------------------------------------------------ index maps
def brev(i: int, bits: int) -> int:
    return int(lib().orc_brev(i, bits))


def slot_freq(order: str, n: int) -> np.ndarray:
    """Frequency held by each output slot (orc_slot_map: orc_slot_freq for every slot)."""
    out = np.empty(n, np.int64)
    lib().orc_slot_map(ORDERS[order], n, out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def unordered_q(i: int) -> int:
    return int(lib().orc_unordered_q(i))


def unordered_q_network(i: int) -> int:
    return int(lib().orc_unordered_q_network(i))


# ------------------------------------------------------------------ FP64 transforms
def dft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    n = len(x)
    xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    yr, yi = np.empty(n), np.empty(n)
    lib().orc_dft(_f64(xr), _f64(xi), _f64(yr), _f64(yi), n, int(inverse))
    return yr + 1j * yi


def fft64(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    n = len(x)
    xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    yr, yi = np.empty(n), np.empty(n)
    lib().orc_fft(_f64(xr), _f64(xi), _f64(yr), _f64(yi), n, int(inverse))
    return yr + 1j * yi


def fft_unit(order: str, inverse: bool, normalized: bool, n: int, src: np.ndarray, src_pack: int,
             dst_pack: int, src_size: int | None = None):
    """Reference fft_unit semantics in FP64.  src: float32 packed buffer.
    Returns (dst float32 packed buffer, FP64 values per output slot as complex128)."""
    src = np.ascontiguousarray(src, np.float32)
    dp = dst_pack if inverse else spectrum_pack(order, dst_pack)
    dst = np.zeros(real_size(n, dp), np.float32)
    outr, outi = np.empty(n), np.empty(n)
    lib().orc_fft_unit(ORDERS[order], int(inverse), int(normalized), n, _f32(dst), dst_pack, _f32(src),
                       src_pack, n if src_size is None else src_size, _f64(outr), _f64(outi))
    return dst, outr + 1j * outi


def xcorr_naive(f: np.ndarray, g: np.ndarray) -> np.ndarray:
    f = np.asarray(f, np.complex128)
    g = np.asarray(g, np.complex128)
    fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
    gr, gi = np.ascontiguousarray(g.real), np.ascontiguousarray(g.imag)
    outr, outi = np.empty(len(f)), np.empty(len(f))
    lib().orc_xcorr_naive(_f64(fr), _f64(fi), len(f), _f64(gr), _f64(gi), len(g), _f64(outr), _f64(outi))
    return outr + 1j * outi


def fxcorr_overlap_save(f: np.ndarray, g: np.ndarray, fft_size: int, pack: int = 8):
    f = np.asarray(f, np.complex128)
    g = np.asarray(g, np.complex128)
    fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
    gr, gi = np.ascontiguousarray(g.real), np.ascontiguousarray(g.imag)
    outr, outi = np.zeros(len(f)), np.zeros(len(f))
    nb = lib().orc_fxcorr_overlap_save(_f64(fr), _f64(fi), len(f), _f64(gr), _f64(gi), len(g), fft_size,
                                       pack, _f64(outr), _f64(outi))
    return outr + 1j * outi, int(nb)


# ------------------------------------------------------------------ FP32 host-core port
class PortPlan:
    def __init__(self, n: int):
        self.n = n
        self._p = lib().orc_port_plan_create(n)
        if not self._p:
            raise ValueError(f"bad size {n}")

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_port_plan_destroy(self._p)
            self._p = None

    def fft(self, z: np.ndarray, order: str = "normal", inverse: bool = False, normalized: bool = True):
        re = np.ascontiguousarray(np.real(z), np.float32)
        im = np.ascontiguousarray(np.imag(z), np.float32)
        lib().orc_port_fft(self._p, ORDERS[order], int(inverse), int(normalized), _f32(re), _f32(im))
        return re + 1j * im

    def batch(self, re: np.ndarray, im: np.ndarray, dist: int, batch: int, order: str = "normal",
              inverse: bool = False, normalized: bool = True, threads: int = 1) -> None:
        lib().orc_port_fft_batch(self._p, ORDERS[order], int(inverse), int(normalized), _f32(re), _f32(im),
                                 dist, batch, threads)


def port_fxcorr(f: np.ndarray, g: np.ndarray, fft_size: int, threads: int = 1):
    fr = np.ascontiguousarray(np.real(f), np.float32)
    fi = np.ascontiguousarray(np.imag(f), np.float32)
    gr = np.ascontiguousarray(np.real(g), np.float32)
    gi = np.ascontiguousarray(np.imag(g), np.float32)
    outr = np.zeros(len(f), np.float32)
    outi = np.zeros(len(f), np.float32)
    nb = lib().orc_port_fxcorr(_f32(fr), _f32(fi), len(f), _f32(gr), _f32(gi), len(g), fft_size, _f32(outr),
                               _f32(outi), threads)
    return outr + 1j * outi, int(nb)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


# ---------------------------------------------------------------------------- vector expressions
def cdiv_ref(a, b):
    """simd_common.hpp:296-309 complex division: (a * conj(b)) / (br^2 + bi^2), in FP64."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = b.real * b.real + b.imag * b.imag
    return ((a.real * b.real + a.imag * b.imag) + 1j * (a.imag * b.real - a.real * b.imag)) / den


def vexpr_run(prog, srcs):
    """FP64 restatement of the four-register machine that pcx_vexpr_eval runs per element
    (include/pcx_gpu.h pcx_vop); srcs are complex arrays.  Real-scalar ops follow
    simd_common.hpp:368-410 (add/sub touch the real part only)."""
    names = ["LOAD", "ADD", "SUB", "MUL", "DIV", "CONJ", "ADDS", "SUBS", "RSUBS", "MULS", "DIVS", "RDIVS",
             "ADDR", "SUBR", "RSUBR", "MULR", "DIVR", "RDIVR"]
    n = len(srcs[0]) if srcs else 0
    r = [np.zeros(n, np.complex128) for _ in range(4)]
    for op, d, a, b, sre, sim in prog:
        name = names[op]
        x, y, s = r[a] if name != "LOAD" else None, r[b], complex(sre, sim)
        if name == "LOAD":
            z = np.asarray(srcs[a], np.complex128)
        elif name == "ADD":
            z = x + y
        elif name == "SUB":
            z = x - y
        elif name == "MUL":
            z = x * y
        elif name == "DIV":
            z = cdiv_ref(x, y)
        elif name == "CONJ":
            z = np.conj(x)
        elif name == "ADDS":
            z = x + s
        elif name == "SUBS":
            z = x - s
        elif name == "RSUBS":
            z = s - x
        elif name == "MULS":
            z = x * s
        elif name == "DIVS":
            z = cdiv_ref(x, s)
        elif name == "RDIVS":
            z = cdiv_ref(s, x)
        elif name == "ADDR":
            z = (x.real + sre) + 1j * x.imag
        elif name == "SUBR":
            z = (x.real - sre) + 1j * x.imag
        elif name == "RSUBR":
            z = (sre - x.real) - 1j * x.imag
        elif name == "MULR":
            z = x * sre
        elif name == "DIVR":
            z = x / sre
        else:  # RDIVR
            z = cdiv_ref(sre, x)
        r[d] = z
    return r[0]
