"""CPU checks of the C ABI boundary: the library loads, exports every symbol include/pcx_gpu.h
declares, and validates arguments with the reference's messages before touching the GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pcx_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcx_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding():
    from packed_amd import _lib

    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from packed_amd import _lib

    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def _create(n, sub=2048, order=0):
    from packed_amd import _lib

    h = C.c_void_p()
    rc = _lib.lib().pcx_plan_create(n, sub, order, C.byref(h))
    return rc, _lib.lib().pcx_last_error().decode()


@pytest.mark.parametrize("n", [0, 1, 3, 100, 1000, -8])
def test_invalid_size_message(n):
    from packed_amd import _lib

    rc, msg = _create(n)
    assert rc == _lib.ERR_INVALID_SIZE
    assert msg == f"fft_size (which is  {n}) is not an integer power of two"  # fft.hpp:1806-1812


@pytest.mark.parametrize("sub,tail", [(4, "is smaller than minimum (which is 8)"),
                                      (96, "is not an integer power of two")])
def test_invalid_sub_size_message(sub, tail):
    from packed_amd import _lib

    rc, msg = _create(1024, sub)
    assert rc == _lib.ERR_INVALID_SUB_SIZE
    assert msg == f"fft_sub_size (which is  {sub}) {tail}"  # fft.hpp:1814-1826


def test_null_arguments():
    from packed_amd import _lib

    L = _lib.lib()
    assert L.pcx_plan_create(1024, 2048, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_fft(None, None, 8, None, 8, 8, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_fxcorr_exec(None, None, 8, None, 8, 8, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_plan_size(None) == 0
    assert L.pcx_version() >= 1


def test_f64_plan_validation():
    """fft_unit<double, Order> checks (fft.hpp:1806-1826 messages); refused orders and sizes fail
    before any device work."""
    from packed_amd import _lib

    L = _lib.lib()
    h = C.c_void_p()
    assert L.pcx_plan_create_f64(1000, 2048, 0, C.byref(h)) == _lib.ERR_INVALID_SIZE
    assert L.pcx_last_error().decode() == "fft_size (which is  1000) is not an integer power of two"
    assert L.pcx_plan_create_f64(1024, 4, 0, C.byref(h)) == _lib.ERR_INVALID_SUB_SIZE
    assert L.pcx_plan_create_f64(1024, 2048, 2, C.byref(h)) == _lib.ERR_UNSUPPORTED  # no double unordered
    assert L.pcx_plan_create_f64(1 << 24, 2048, 0, C.byref(h)) == _lib.ERR_UNSUPPORTED
    assert L.pcx_plan_create_f64(1024, 2048, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_fft_f64(None, None, 4, None, 4, 4, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_ifft_f64(None, 1, None, 4, None, 4, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_fft_batch_f64(None, 0, 1, None, 4, 8, None, 4, 8, 4, 2, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.pcx_plan_element_size(None) == 0
