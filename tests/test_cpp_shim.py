"""Runs the C++ drop-in test (tests/cpp/test_fft_shim.cpp, built by __graft_entry__.build()):
the pcx_gpu::fft_unit / fxcorr_unit classes with the reference's template signatures, checked
against the FP64 oracle like the reference's own tests/test_fft.cpp and tests/test_fxcorr.cpp."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_fft_shim")


def test_shim_header_compiles_standalone(tmp_path):
    """The header needs only pcx_gpu.h and a C++20 compiler (no HIP headers)."""
    root = os.path.dirname(HERE)
    src = tmp_path / "t.cpp"
    src.write_text('#include "pcx_gpu/fft.hpp"\nint main(){ pcx_gpu::fft_unit<float> (*f)(size_t) = nullptr; (void)f; }\n')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{root}/include", f"-I{root}/packed_amd/include",
                    str(src)], check=True)


def _syntax(tmp_path, body):
    root = os.path.dirname(HERE)
    src = tmp_path / "t.cpp"
    src.write_text('#include "pcx_gpu/fft.hpp"\nint main(){ ' + body + ' }\n')
    return subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{root}/include", f"-I{root}/packed_amd/include",
                           str(src)], capture_output=True, text=True)


def test_double_unit_instantiates(tmp_path):
    """fft_unit<double, Order> with the reference's default pack 4 (types.hpp:70-72)."""
    body = ("pcx_gpu::fft_unit<double, pcx_gpu::fft_order::bit_reversed>* u = nullptr; "
            "pcx_gpu::vector<double>* v = nullptr; if (u) { (*u)(*v); u->ifft(*v); } "
            "static_assert(pcx_gpu::vector<double>::pack_size == 4 && pcx_gpu::vector<float>::pack_size == 8);")
    res = _syntax(tmp_path, body)
    assert res.returncode == 0, res.stderr


def test_double_unordered_is_a_compile_error(tmp_path):
    res = _syntax(tmp_path, "pcx_gpu::fft_unit<double, pcx_gpu::fft_order::unordered> u(1024);")
    assert res.returncode != 0 and "unordered order is defined for float only" in res.stderr


def test_mixed_element_types_are_a_compile_error(tmp_path):
    res = _syntax(tmp_path, "pcx_gpu::fft_unit<float>* u = nullptr; pcx_gpu::vector<double>* v = nullptr; (*u)(*v);")
    assert res.returncode != 0


@pytest.mark.parametrize("body", [
    # tests/test_fft.cpp:307, 463-464, 719-720 instantiation forms
    "auto u = pcx_gpu::fft_unit<float, pcx_gpu::fft_order::normal>(1024, 256); (void)u;",
    "pcx_gpu::fft_unit_par<float, pcx_gpu::fft_order::normal, pcx_gpu::device_allocator<float>, 4>* p = nullptr; (void)p;",
    # all seven parameters (fft.hpp:565-571) and the fixed-SubSize constructor (fft.hpp:594-598)
    "using U = pcx_gpu::fft_unit<float, pcx_gpu::fft_order::bit_reversed, std::allocator<float>, 2048, 8, "
    "pcx_gpu::strategy8, pcx_gpu::strategy48>; U* u = nullptr; if (u) { U v(4096); U w(4096, std::allocator<float>{}); }",
    "using U = pcx_gpu::fft_unit<double, pcx_gpu::fft_order::normal, std::allocator<double>, pcx_gpu::dynamic_size, 4, "
    "pcx_gpu::strategy42, pcx_gpu::strategy42>; U* u = nullptr; if (u) { U v(4096, 512); }",
    "pcx_gpu::fft_unit_par<double>* p = nullptr; (void)p;",
    # fxcorr_unit's policy slots (fxcorr.hpp:44-48)
    "using X = pcx_gpu::fxcorr_unit<float, pcx_gpu::device_allocator<float>, "
    "pcx_gpu::fft_unit<float, pcx_gpu::fft_order::normal>, pcx_gpu::detail_::pseudo_vector_factory<float>>; "
    "X* x = nullptr; (void)x;",
])
def test_reference_template_forms_compile(tmp_path, body):
    res = _syntax(tmp_path, body)
    assert res.returncode == 0, res.stderr


@pytest.mark.parametrize("body", [
    # SubSize must be a power of two >= the default pack size (fft.hpp:572-574)
    "pcx_gpu::fft_unit<float, pcx_gpu::fft_order::normal, std::allocator<float>, 12>* u = nullptr; (void)u;",
    "pcx_gpu::fft_unit<float, pcx_gpu::fft_order::normal, std::allocator<float>, 4>* u = nullptr; (void)u;",
    # the one-argument constructor belongs to a fixed SubSize, the sub_size argument to a dynamic one
    "pcx_gpu::fft_unit<float, pcx_gpu::fft_order::normal, std::allocator<float>, 2048> u(1024, 256);",
])
def test_reference_template_constraints(tmp_path, body):
    assert _syntax(tmp_path, body).returncode != 0


@pytest.mark.gpu
def test_cpp_drop_in():
    assert os.path.exists(BIN), "build the C++ test first: python __graft_entry__.py"
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-2000:]
