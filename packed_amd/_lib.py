"""ctypes binding of libpcx_gpu.so (the C ABI declared in include/pcx_gpu.h).

The shared library is built in-tree by ``packed_amd.build.build()``.  Loading fails loudly when
it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcx_gpu.so")

# pcx_status (include/pcx_gpu.h)
SUCCESS = 0
ERR_INVALID_SIZE = 1
ERR_INVALID_SUB_SIZE = 2
ERR_SIZE_MISMATCH = 3
ERR_INPLACE_REPACK = 4
ERR_INVALID_PACK = 5
ERR_INVALID_ARGUMENT = 6
ERR_UNSUPPORTED = 7
ERR_HIP = 8
ERR_OUT_OF_MEMORY = 9

ORDER = {"normal": 0, "bit_reversed": 1, "unordered": 2}
PRECISION = {"fp32": 0, "bf16x3": 1, "fp16x2": 2}

class Job(C.Structure):
    """pcx_job (include/pcx_gpu.h)."""

    _fields_ = [("log2n", C.c_int64), ("inverse", C.c_int32), ("reserved", C.c_int32),
                ("src_offset", C.c_int64), ("dst_offset", C.c_int64)]


class VSource(C.Structure):
    """pcx_vsource (include/pcx_gpu.h)."""

    _fields_ = [("data", C.c_void_p), ("pack", C.c_int64)]


class VInstr(C.Structure):
    """pcx_vinstr (include/pcx_gpu.h)."""

    _fields_ = [("op", C.c_int32), ("d", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("s_re", C.c_float), ("s_im", C.c_float)]


class VSource64(C.Structure):
    """pcx_vsource_f64 (include/pcx_gpu.h)."""

    _fields_ = [("data", C.c_void_p), ("pack", C.c_int64)]


class VInstr64(C.Structure):
    """pcx_vinstr_f64 (include/pcx_gpu.h)."""

    _fields_ = [("op", C.c_int32), ("d", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("s_re", C.c_double), ("s_im", C.c_double)]


# pcx_vop (include/pcx_gpu.h)
VOP = {name: i for i, name in enumerate(
    ["LOAD", "ADD", "SUB", "MUL", "DIV", "CONJ", "ADDS", "SUBS", "RSUBS", "MULS", "DIVS", "RDIVS",
     "ADDR", "SUBR", "RSUBR", "MULR", "DIVR", "RDIVR"])}
VEXPR_MAX_PROG, VEXPR_MAX_SRC, VEXPR_REGS = 64, 16, 4

# Every symbol include/pcx_gpu.h declares, with (restype, argtypes).
_vp, _i64, _int, _sz, _u64, _dbl = C.c_void_p, C.c_int64, C.c_int, C.c_size_t, C.c_uint64, C.c_double
SIGNATURES = {
    "pcx_version": (_int, []),
    "pcx_last_error": (C.c_char_p, []),
    "pcx_plan_create": (_int, [_i64, _i64, _int, C.POINTER(_vp)]),
    "pcx_plan_destroy": (_int, [_vp]),
    "pcx_plan_size": (_i64, [_vp]),
    "pcx_plan_order": (_int, [_vp]),
    "pcx_plan_set_precision": (_int, [_vp, _int]),
    "pcx_plan_precision": (_int, [_vp]),
    "pcx_plan_set_shape": (_int, [_vp, _int, _int, _int]),
    "pcx_plan_shape": (_int, [_vp, C.POINTER(_int), C.POINTER(_int), C.POINTER(_int)]),
    "pcx_plan_set_grid": (_int, [_vp, _i64, _i64]),
    "pcx_plan_set_chunk": (_int, [_vp, _i64]),
    "pcx_plan_set_one_pass": (_int, [_vp, _int]),
    "pcx_plan_set_three_pass": (_int, [_vp, _int]),
    "pcx_fft": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_ifft": (_int, [_vp, _int, _vp, _i64, _vp, _i64, _vp]),
    "pcx_fft_batch": (_int, [_vp, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "pcx_fft_batch_noalloc": (_int, [_vp, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "pcx_plan_create_f64": (_int, [_i64, _i64, _int, C.POINTER(_vp)]),
    "pcx_plan_element_size": (_int, [_vp]),
    "pcx_fft_f64": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_ifft_f64": (_int, [_vp, _int, _vp, _i64, _vp, _i64, _vp]),
    "pcx_fft_batch_f64": (_int, [_vp, _int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "pcx_fxcorr_create": (_int, [_vp, _i64, _i64, _i64, _vp, C.POINTER(_vp)]),
    "pcx_fxcorr_destroy": (_int, [_vp]),
    "pcx_fxcorr_exec": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_fxcorr_step": (_i64, [_vp, _i64]),
    "pcx_fxcorr_exec_batch": (_int, [_vp, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "pcx_fxcorr_set_precision": (_int, [_vp, _int]),
    "pcx_fxcorr_workspace_min": (_i64, [_vp]),
    "pcx_fxcorr_exec_batch_ws": (_int, [_vp, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp, _i64, _vp]),
    "pcx_queue_create": (_int, [C.POINTER(Job), _i64, _int, C.POINTER(_vp)]),
    "pcx_queue_destroy": (_int, [_vp]),
    "pcx_queue_exec": (_int, [_vp, _vp, _vp, _vp]),
    "pcx_queue_autotune": (_int, [_vp, _vp, _vp, _vp, C.c_char_p]),
    "pcx_queue_capture": (_int, [_vp, _vp, _vp, _vp]),
    "pcx_queue_launch": (_int, [_vp, _vp]),
    "pcx_queue_work": (_int, [_vp, C.POINTER(_dbl), C.POINTER(_dbl)]),
    "pcx_malloc": (_int, [C.POINTER(_vp), _sz]),
    "pcx_free": (_int, [_vp]),
    "pcx_memcpy_htod": (_int, [_vp, _vp, _sz]),
    "pcx_memcpy_dtoh": (_int, [_vp, _vp, _sz]),
    "pcx_memcpy_dtod": (_int, [_vp, _vp, _sz]),
    "pcx_memset": (_int, [_vp, _int, _sz]),
    "pcx_synchronize": (_int, [_vp]),
    "pcx_generate": (_int, [_vp, _i64, _i64, _u64, _int, _dbl, _vp]),
    "pcx_generate_f64": (_int, [_vp, _i64, _i64, _u64, _int, _dbl, _vp]),
    "pcx_vexpr_eval": (_int, [C.POINTER(VSource), C.c_int32, C.POINTER(VInstr), C.c_int32, _vp, _i64, _i64, _vp]),
    "pcx_repack": (_int, [_vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_vexpr_eval_f64": (_int, [C.POINTER(VSource64), C.c_int32, C.POINTER(VInstr64), C.c_int32, _vp, _i64, _i64,
                                  _vp]),
    "pcx_repack_f64": (_int, [_vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_fft_across": (_int, [_vp, _int, _int, _vp, _i64, _vp, _i64, _i64, _vp]),
    "pcx_fft_across_f64": (_int, [_vp, _int, _int, _vp, _i64, _vp, _i64, _i64, _vp]),
}

_lib = None


class PcxError(ValueError):
    """A non-zero pcx_status; the message is the library's (reference-worded) error text."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run packed_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SUCCESS:
        raise PcxError(rc, lib().pcx_last_error().decode())
