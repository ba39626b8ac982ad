// vector_arithm.hpp -- device-side vector expressions for pcx_gpu::vector (drop-in for the
// reference's inc/vector_arithm.hpp operators: binary + - * / of two vectors of any pack sizes,
// + - * / with a float or std::complex<float> scalar on either side, conj, store).
//
// Expressions are lazy trees of small value types.  store(expr, dst) flattens a tree into a
// program for pcx_vexpr_eval (include/pcx_gpu.h): a four-register machine that one fused kernel
// runs per element, so a whole pointwise pipeline reads each operand once and writes once.
// Registers are allocated by Sethi-Ullman numbering (the operand needing more registers is
// evaluated first); trees that need more than four registers throw std::length_error.
#pragma once
#include "fft.hpp"

#include <array>
#include <complex>
#include <stdexcept>

namespace pcx_gpu {
namespace vx {

struct leaf {
    const float* data;
    int64_t      pack;
    uZ           n;
    uZ           size() const { return n; }
    int          need() const { return 1; }
};
template<int Op, typename L, typename R> // Op: PCX_VOP_ADD .. PCX_VOP_DIV
struct bin {
    L   l;
    R   r;
    uZ  size() const { return l.size(); }
    int need() const {
        int a = l.need(), b = r.need();
        return a == b ? a + 1 : (a > b ? a : b);
    }
};
template<int Op, typename E> // Op: one of the scalar pcx_vop codes
struct sca {
    E                   e;
    std::complex<float> s;
    uZ                  size() const { return e.size(); }
    int                 need() const { return e.need(); }
};
template<typename E>
struct cj {
    E   e;
    uZ  size() const { return e.size(); }
    int need() const { return e.need(); }
};

template<typename T>
struct is_node : std::false_type {};
template<>
struct is_node<leaf> : std::true_type {};
template<int Op, typename L, typename R>
struct is_node<bin<Op, L, R>> : std::true_type {};
template<int Op, typename E>
struct is_node<sca<Op, E>> : std::true_type {};
template<typename E>
struct is_node<cj<E>> : std::true_type {};

// any pcx_gpu::vector<float, P> or expression node (expressions are evaluated in FP32)
template<typename X>
concept operand = is_node<std::remove_cvref_t<X>>::value ||
                  (device_complex_vector<std::remove_cvref_t<X>> &&
                   std::same_as<typename std::remove_cvref_t<X>::real_type, float>);
template<typename S>
concept scalar = std::same_as<S, float> || std::same_as<S, std::complex<float>>;

template<operand X>
auto node(const X& x) {
    if constexpr (is_node<X>::value)
        return x;
    else
        return leaf{x.data(), (int64_t)X::pack_size, x.size()};
}

struct program {
    std::array<pcx_vinstr, PCX_VEXPR_MAX_PROG>  code{};
    std::array<pcx_vsource, PCX_VEXPR_MAX_SRC> src{};
    int                                         ncode = 0, nsrc = 0;

    void emit(int op, int d, int a, int b = 0, std::complex<float> s = {}) {
        if (ncode == PCX_VEXPR_MAX_PROG)
            throw std::length_error("vector expression exceeds 64 instructions");
        code[ncode++] = pcx_vinstr{op, d, a, b, s.real(), s.imag()};
    }
    int source(const leaf& v) {
        for (int k = 0; k < nsrc; ++k)
            if (src[k].data == v.data && src[k].pack == v.pack)
                return k;
        if (nsrc == PCX_VEXPR_MAX_SRC)
            throw std::length_error("vector expression reads more than 16 vectors");
        src[nsrc] = pcx_vsource{v.data, v.pack};
        return nsrc++;
    }
    void gen(const leaf& v, int r) { emit(PCX_VOP_LOAD, r, source(v)); }
    template<int Op, typename L, typename R>
    void gen(const bin<Op, L, R>& e, int r) {
        if (e.l.need() >= e.r.need()) {
            gen(e.l, r);
            gen(e.r, r + 1);
            emit(Op, r, r, r + 1);
        } else {
            gen(e.r, r);
            gen(e.l, r + 1);
            emit(Op, r, r + 1, r);
        }
    }
    template<int Op, typename E>
    void gen(const sca<Op, E>& e, int r) {
        gen(e.e, r);
        emit(Op, r, r, 0, e.s);
    }
    template<typename E>
    void gen(const cj<E>& e, int r) {
        gen(e.e, r);
        emit(PCX_VOP_CONJ, r, r);
    }
};

template<int Op, operand A, operand B>
auto make_bin(const A& a, const B& b) {
    auto l = node(a);
    auto r = node(b);
    if (l.size() != r.size())
        throw std::invalid_argument("vector sizes (which are " + std::to_string(l.size()) + " and " +
                                    std::to_string(r.size()) + ") differ");
    return bin<Op, decltype(l), decltype(r)>{l, r};
}
// scalar on the right (Rop) or left (Lop), real (float) or complex forms
template<int CxR, int CxL, int ReR, int ReL, bool Left, operand A, scalar S>
auto make_sca(const A& a, S s) {
    auto e = node(a);
    if constexpr (std::same_as<S, float>)
        return sca<Left ? ReL : ReR, decltype(e)>{e, std::complex<float>(s, 0.f)};
    else
        return sca<Left ? CxL : CxR, decltype(e)>{e, s};
}

} // namespace vx

template<int Op, typename L, typename R>
struct is_vexpr<vx::bin<Op, L, R>> : std::true_type {};
template<int Op, typename E>
struct is_vexpr<vx::sca<Op, E>> : std::true_type {};
template<typename E>
struct is_vexpr<vx::cj<E>> : std::true_type {};

// vector_arithm.hpp:453-463 (vector op vector)
template<vx::operand A, vx::operand B>
auto operator+(const A& a, const B& b) { return vx::make_bin<PCX_VOP_ADD>(a, b); }
template<vx::operand A, vx::operand B>
auto operator-(const A& a, const B& b) { return vx::make_bin<PCX_VOP_SUB>(a, b); }
template<vx::operand A, vx::operand B>
auto operator*(const A& a, const B& b) { return vx::make_bin<PCX_VOP_MUL>(a, b); }
template<vx::operand A, vx::operand B>
auto operator/(const A& a, const B& b) { return vx::make_bin<PCX_VOP_DIV>(a, b); }
// vector_arithm.hpp:465-492 (with a scalar on either side)
template<vx::operand A, vx::scalar S>
auto operator+(const A& a, S s) { return vx::make_sca<PCX_VOP_ADDS, 0, PCX_VOP_ADDR, 0, false>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator+(S s, const A& a) { return vx::make_sca<0, PCX_VOP_ADDS, 0, PCX_VOP_ADDR, true>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator-(const A& a, S s) { return vx::make_sca<PCX_VOP_SUBS, 0, PCX_VOP_SUBR, 0, false>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator-(S s, const A& a) { return vx::make_sca<0, PCX_VOP_RSUBS, 0, PCX_VOP_RSUBR, true>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator*(const A& a, S s) { return vx::make_sca<PCX_VOP_MULS, 0, PCX_VOP_MULR, 0, false>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator*(S s, const A& a) { return vx::make_sca<0, PCX_VOP_MULS, 0, PCX_VOP_MULR, true>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator/(const A& a, S s) { return vx::make_sca<PCX_VOP_DIVS, 0, PCX_VOP_DIVR, 0, false>(a, s); }
template<vx::operand A, vx::scalar S>
auto operator/(S s, const A& a) { return vx::make_sca<0, PCX_VOP_RDIVS, 0, PCX_VOP_RDIVR, true>(a, s); }
// vector_arithm.hpp:495
template<vx::operand A>
auto conj(const A& a) {
    auto e = vx::node(a);
    return vx::cj<decltype(e)>{e};
}

// vector_arithm.hpp:405-447: evaluate an expression (or copy / repack a vector) into dest
template<typename E, typename V>
void store(const E& expr, V& dest) {
    auto e = vx::node(expr);
    if (e.size() != dest.size())
        throw std::invalid_argument("expression size (which is " + std::to_string(e.size()) +
                                    ") differs from destination size (which is " + std::to_string(dest.size()) + ")");
    if (e.need() > PCX_VEXPR_REGS)
        throw std::length_error("vector expression needs more than four registers");
    vx::program p;
    p.gen(e, 0);
    throw_on(pcx_vexpr_eval(p.src.data(), p.nsrc, p.code.data(), p.ncode, dest.data(), (int64_t)V::pack_size,
                            (int64_t)dest.size(), nullptr));
}

} // namespace pcx_gpu
