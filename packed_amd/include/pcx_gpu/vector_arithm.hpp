// vector_arithm.hpp -- device-side vector expressions for pcx_gpu::vector (drop-in for the
// reference's inc/vector_arithm.hpp operators: binary + - * / of two vectors of any pack sizes,
// + - * / with a T or std::complex<T> scalar on either side, conj, store; T = float or double,
// one element type per expression).
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

template<typename T>
struct leaf {
    using real_type = T;
    const T* data;
    int64_t  pack;
    uZ       n;
    uZ       size() const { return n; }
    int      need() const { return 1; }
};
template<int Op, typename L, typename R> // Op: PCX_VOP_ADD .. PCX_VOP_DIV
struct bin {
    using real_type = typename L::real_type;
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
    using real_type = typename E::real_type;
    E                          e;
    std::complex<real_type>    s;
    uZ                         size() const { return e.size(); }
    int                        need() const { return e.need(); }
};
template<typename E>
struct cj {
    using real_type = typename E::real_type;
    E   e;
    uZ  size() const { return e.size(); }
    int need() const { return e.need(); }
};

template<typename T>
struct is_node : std::false_type {};
template<typename T>
struct is_node<leaf<T>> : std::true_type {};
template<int Op, typename L, typename R>
struct is_node<bin<Op, L, R>> : std::true_type {};
template<int Op, typename E>
struct is_node<sca<Op, E>> : std::true_type {};
template<typename E>
struct is_node<cj<E>> : std::true_type {};

// any pcx_gpu::vector<T, P> (T = float or double) or expression node
template<typename X>
concept operand = is_node<std::remove_cvref_t<X>>::value ||
                  (device_complex_vector<std::remove_cvref_t<X>> &&
                   gpu_real<typename std::remove_cvref_t<X>::real_type>);
template<typename X>
using real_of = typename std::remove_cvref_t<X>::real_type;
// a scalar of the expression's element type: T or std::complex<T>
template<typename S, typename T>
concept scalar_of = std::same_as<S, T> || std::same_as<S, std::complex<T>>;

template<operand X>
auto node(const X& x) {
    if constexpr (is_node<X>::value)
        return x;
    else
        return leaf<real_of<X>>{x.data(), (int64_t)X::pack_size, x.size()};
}

template<typename T>
struct program {
    static constexpr bool f64 = std::same_as<T, double>;
    using instr               = std::conditional_t<f64, pcx_vinstr_f64, pcx_vinstr>;
    using source_t            = std::conditional_t<f64, pcx_vsource_f64, pcx_vsource>;
    std::array<instr, PCX_VEXPR_MAX_PROG>     code{};
    std::array<source_t, PCX_VEXPR_MAX_SRC> src{};
    int                                       ncode = 0, nsrc = 0;

    void emit(int op, int d, int a, int b = 0, std::complex<T> s = {}) {
        if (ncode == PCX_VEXPR_MAX_PROG)
            throw std::length_error("vector expression exceeds 64 instructions");
        code[ncode++] = instr{op, d, a, b, s.real(), s.imag()};
    }
    int source(const leaf<T>& v) {
        for (int k = 0; k < nsrc; ++k)
            if (src[k].data == v.data && src[k].pack == v.pack)
                return k;
        if (nsrc == PCX_VEXPR_MAX_SRC)
            throw std::length_error("vector expression reads more than 16 vectors");
        src[nsrc] = source_t{v.data, v.pack};
        return nsrc++;
    }
    void gen(const leaf<T>& v, int r) { emit(PCX_VOP_LOAD, r, source(v)); }
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
    static_assert(std::same_as<real_of<A>, real_of<B>>, "an expression has one element type (float or double)");
    auto l = node(a);
    auto r = node(b);
    if (l.size() != r.size())
        throw std::invalid_argument("vector sizes (which are " + std::to_string(l.size()) + " and " +
                                    std::to_string(r.size()) + ") differ");
    return bin<Op, decltype(l), decltype(r)>{l, r};
}
// scalar on the right (Rop) or left (Lop), real (T) or complex forms
template<int CxR, int CxL, int ReR, int ReL, bool Left, operand A, typename S>
auto make_sca(const A& a, S s) {
    using T = real_of<A>;
    auto e  = node(a);
    if constexpr (std::same_as<S, T>)
        return sca<Left ? ReL : ReR, decltype(e)>{e, std::complex<T>(s, T(0))};
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
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator+(const A& a, S s) { return vx::make_sca<PCX_VOP_ADDS, 0, PCX_VOP_ADDR, 0, false>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator+(S s, const A& a) { return vx::make_sca<0, PCX_VOP_ADDS, 0, PCX_VOP_ADDR, true>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator-(const A& a, S s) { return vx::make_sca<PCX_VOP_SUBS, 0, PCX_VOP_SUBR, 0, false>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator-(S s, const A& a) { return vx::make_sca<0, PCX_VOP_RSUBS, 0, PCX_VOP_RSUBR, true>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator*(const A& a, S s) { return vx::make_sca<PCX_VOP_MULS, 0, PCX_VOP_MULR, 0, false>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator*(S s, const A& a) { return vx::make_sca<0, PCX_VOP_MULS, 0, PCX_VOP_MULR, true>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
auto operator/(const A& a, S s) { return vx::make_sca<PCX_VOP_DIVS, 0, PCX_VOP_DIVR, 0, false>(a, s); }
template<vx::operand A, typename S>
    requires vx::scalar_of<S, vx::real_of<A>>
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
    using T = vx::real_of<E>;
    static_assert(std::same_as<T, typename V::real_type>, "an expression and its destination share the element type");
    auto e  = vx::node(expr);
    if (e.size() != dest.size())
        throw std::invalid_argument("expression size (which is " + std::to_string(e.size()) +
                                    ") differs from destination size (which is " + std::to_string(dest.size()) + ")");
    if (e.need() > PCX_VEXPR_REGS)
        throw std::length_error("vector expression needs more than four registers");
    vx::program<T> p;
    p.gen(e, 0);
    if constexpr (std::same_as<T, double>)
        throw_on(pcx_vexpr_eval_f64(p.src.data(), p.nsrc, p.code.data(), p.ncode, dest.data(), (int64_t)V::pack_size,
                                    (int64_t)dest.size(), nullptr));
    else
        throw_on(pcx_vexpr_eval(p.src.data(), p.nsrc, p.code.data(), p.ncode, dest.data(), (int64_t)V::pack_size,
                                (int64_t)dest.size(), nullptr));
}

} // namespace pcx_gpu
