// vexpr_kernels.hpp -- fused element-wise evaluation of packed complex vector expressions.
//
// A program of up to 64 instructions runs per element on four complex registers (held in VGPRs;
// register numbers are wave-uniform, so operand selection is a short chain of uniform selects).
// Every source is read once per element and the destination written once: the kernel moves
// 8 B per source element plus 8 B per output element and is bound by HBM bandwidth.
// Arithmetic follows simd_common.hpp:296-410 of the reference (complex division as
// (a conj(b)) / |b|^2 with fused multiply-adds; real scalars act on the real part for + and -).
#pragma once
#include "fft_kernels.hpp"

namespace pcxk {

// T = float or double: the element type of every vector and scalar of the program.
template<class T>
struct VInstrT {
    int32_t op, d, a, b;
    T       s_re, s_im;
};
template<class T>
struct VSourceT {
    const T* data;
    int64_t  lp; // log2 of the pack size
};
constexpr int kVMaxProg = 64, kVMaxSrc = 16, kVRegs = 4;
template<class T>
struct VExprArgsT {
    VSourceT<T> src[kVMaxSrc];
    VInstrT<T>  prog[kVMaxProg];
    int32_t     nprog;
    T*          dst;
    int         dst_lp;
    int64_t     n;
};

enum : int32_t {
    VOP_LOAD = 0, VOP_ADD, VOP_SUB, VOP_MUL, VOP_DIV, VOP_CONJ, VOP_ADDS, VOP_SUBS, VOP_RSUBS, VOP_MULS, VOP_DIVS,
    VOP_RDIVS, VOP_ADDR, VOP_SUBR, VOP_RSUBR, VOP_MULR, VOP_DIVR, VOP_RDIVR, VOP_COUNT
};

// simd_common.hpp:296-309: (a conj(b)) / (br^2 + bi^2)
template<class T>
__device__ __forceinline__ cx<T> vdiv(cx<T> a, cx<T> b) {
    T den = fmat(b.y, b.y, b.x * b.x);
    T re  = fmat(a.y, b.y, a.x * b.x);
    T im  = fmat(a.y, b.x, -(a.x * b.y));
    return {re / den, im / den};
}
// simd_common.hpp:372-383: real s / b
template<class T>
__device__ __forceinline__ cx<T> vrdiv_real(T s, cx<T> b) {
    T den = fmat(b.y, b.y, b.x * b.x);
    return {(s * b.x) / den, -(s * b.y) / den};
}

template<class T>
__device__ __forceinline__ cx<T> vsel(const cx<T> (&r)[kVRegs], int i) {
    cx<T> v = r[0];
    v    = i == 1 ? r[1] : v;
    v    = i == 2 ? r[2] : v;
    v    = i == 3 ? r[3] : v;
    return v;
}

template<class T>
__global__ __launch_bounds__(256) void k_vexpr(VExprArgsT<T> p) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += stride) {
        cx<T> r[kVRegs] = {};
        for (int k = 0; k < p.nprog; ++k) {
            const VInstrT<T> in = p.prog[k]; // uniform: kernel arguments
            const cx<T>      x  = vsel(r, in.a);
            const cx<T>      y  = vsel(r, in.b);
            const cx<T>      s  = {in.s_re, in.s_im};
            cx<T>            z;
            switch (in.op) {
            case VOP_LOAD: {
                const VSourceT<T> src = p.src[in.a];
                const int64_t o   = pidx(i, (int)src.lp);
                z                 = {src.data[o], src.data[o + ((int64_t)1 << src.lp)]};
                break;
            }
            case VOP_ADD: z = cadd(x, y); break;
            case VOP_SUB: z = csub(x, y); break;
            case VOP_MUL: z = cmul(x, y); break;
            case VOP_DIV: z = vdiv(x, y); break;
            case VOP_CONJ: z = cconj(x); break;
            case VOP_ADDS: z = cadd(x, s); break;
            case VOP_SUBS: z = csub(x, s); break;
            case VOP_RSUBS: z = csub(s, x); break;
            case VOP_MULS: z = cmul(x, s); break;
            case VOP_DIVS: z = vdiv(x, s); break;
            case VOP_RDIVS: z = vdiv(s, x); break;
            case VOP_ADDR: z = {x.x + in.s_re, x.y}; break;
            case VOP_SUBR: z = {x.x - in.s_re, x.y}; break;
            case VOP_RSUBR: z = {in.s_re - x.x, -x.y}; break;
            case VOP_MULR: z = {x.x * in.s_re, x.y * in.s_re}; break;
            case VOP_DIVR: z = {x.x / in.s_re, x.y / in.s_re}; break;
            case VOP_RDIVR: z = vrdiv_real(in.s_re, x); break;
            default: z = x; break;
            }
            r[0] = in.d == 0 ? z : r[0];
            r[1] = in.d == 1 ? z : r[1];
            r[2] = in.d == 2 ? z : r[2];
            r[3] = in.d == 3 ? z : r[3];
        }
        const int64_t o                       = pidx(i, p.dst_lp);
        p.dst[o]                              = r[0].x;
        p.dst[o + ((int64_t)1 << p.dst_lp)]   = r[0].y;
    }
}

} // namespace pcxk
