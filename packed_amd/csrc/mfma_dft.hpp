// mfma_dft.hpp -- radix-16 DFT on the gfx950 matrix cores with split-precision operands.
//
// A wavefront's 64 radix-16 butterflies (one per lane, 16 complex values in registers, the layout
// of the FP32 path) are computed as two real 32x32x16 products per operand pair:
//   [Yr; Yi] = [Fr; Fi] * Xr + [-Fi; Fr] * Xi,  F[k][r] = exp(-2*pi*i*k*r/16),
// with the 32 outputs of 32 butterflies in one 32x32 FP32 accumulator
// (v_mfma_f32_32x32x16_bf16).  Operands are split into three BF16 terms: F by truncation, which is
// exact (8 + 8 + 8 bits cover the FP32 significand), the data by round-to-nearest conversion
// (v_cvt_pk_bf16_f32, residual below 2^-24): x = x0 + x1 + x2, F = F0 + F1 + F2, and the six
// products with combined weight >= 2^-16 are accumulated (smallest first) in FP32:
//   F1x1 + F0x2 + F2x0 + F0x1 + F1x0 + F0x0;  the dropped terms are below 2^-24 relative.
//
// Lane layout of v_mfma_f32_32x32x16_{bf16,f16}: A (32x16) lane l holds A[l%32][8*(l/32) + t],
// B (16x32) lane l holds B[8*(l/32) + t][l%32], t < 8; the FP32 accumulator lane l holds
// D[8*(v/4) + 4*(l/32) + v%4][l%32], v < 16.  Two half-wave exchanges (v_permlane32_swap) move
// the butterflies between the one-butterfly-per-lane layout and this one.
#pragma once
#include "dev_common.hpp"

namespace pcxk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  floatx16 __attribute__((ext_vector_type(16)));

enum : int { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP16X2 = 2 };

// The three truncated BF16 terms of x (exact: x == t0 + t1 + t2 for normal FP32 x).
__device__ __forceinline__ void split3(float x, __bf16& t0, __bf16& t1, __bf16& t2) {
    uint32_t u0 = __float_as_uint(x) & 0xFFFF0000u;
    float    r1 = x - __uint_as_float(u0);
    uint32_t u1 = __float_as_uint(r1) & 0xFFFF0000u;
    float    r2 = r1 - __uint_as_float(u1);
    uint32_t u2 = __float_as_uint(r2) & 0xFFFF0000u;
    t0          = __builtin_bit_cast(__bf16, (uint16_t)(u0 >> 16));
    t1          = __builtin_bit_cast(__bf16, (uint16_t)(u1 >> 16));
    t2          = __builtin_bit_cast(__bf16, (uint16_t)(u2 >> 16));
}

typedef __bf16   bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest three-term split of the pair (a, b) into packed BF16 pairs t0, t1, t2 (a in
// the low half) with v_cvt_pk_bf16_f32: a = a0 + a1 + a2 + e, |e| <= 2^-24 |a| (each remainder
// is exact in FP32 and each rounding keeps 8 significant bits).
__device__ __forceinline__ void split3_pk(float a, float b, uint32_t& t0, uint32_t& t1, uint32_t& t2) {
    t0       = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)a, (__bf16)b});
    float ra = a - __uint_as_float(t0 << 16);
    float rb = b - __uint_as_float(t0 & 0xFFFF0000u);
    t1       = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)ra, (__bf16)rb});
    ra -= __uint_as_float(t1 << 16);
    rb -= __uint_as_float(t1 & 0xFFFF0000u);
    t2 = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)ra, (__bf16)rb});
}

// FP16X2: F and the data as hi + lo FP16 pairs (11 + 11 significant bits; the lo term of a
// value is its FP32 remainder rounded once, so x = hi + lo + e with |e| <= 2^-22 |x|), three
// products F_lo*x_hi + F_hi*x_lo + F_hi*x_hi accumulated in FP32 (F_lo*x_lo is below 2^-22).
// FP16 has 5 exponent bits, so each butterfly's 16 inputs are scaled by a power of two 2^e that
// puts their largest magnitude in [2^14, 2^15) (below the FP16 maximum 65504, with the lo terms
// of values down to 2^-10 of that maximum still normal), and the outputs by 2^-e.  Both
// scalings are exact, and the DFT is linear, so the scale drops out.
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

// 2^e (and 2^-e) for the butterfly held in lanes l and l^32 (its column of the B operand):
// m = the largest |x| over both lanes, e = 15 - ceil-exponent(m), clamped to [-126, 126] so
// that 2^e and 2^-e are both normal FP32 numbers.
__device__ __forceinline__ void col_scale(const float* r, const float* i, float& sc, float& inv) {
    float m = 0.f;
#pragma unroll
    for (int t = 0; t < 8; ++t)
        m = fmaxf(m, fmaxf(fabsf(r[t]), fabsf(i[t])));
    auto  w = __builtin_amdgcn_permlane32_swap(__float_as_uint(m), __float_as_uint(m), false, false);
    m       = fmaxf(__uint_as_float(w[0]), __uint_as_float(w[1]));
    int ex  = (int)((__float_as_uint(m) >> 23) & 0xFFu) - 126; // m = f * 2^ex, f in [0.5, 1)
    int e   = 15 - ex;
    e       = e < -126 ? -126 : (e > 126 ? 126 : e);
    sc      = __uint_as_float((uint32_t)(127 + e) << 23);
    inv     = __uint_as_float((uint32_t)(127 - e) << 23);
}

// Per-lane A operands: [Fr; Fi] and [-Fi; Fr] as raw bits: three BF16 terms (BF16X3) or the
// hi and lo FP16 terms in entries 0 and 1 (FP16X2).
struct MfmaDft16 {
    u32x4 a1[3], a2[3];

    template<int PREC = PREC_BF16X3>
    __device__ void init() {
        const int lane = threadIdx.x & 63;
        const int i    = lane & 31;
        const int h    = lane >> 5;
        const int k    = i & 15; // output frequency of this row
        if constexpr (PREC == PREC_FP16X2) {
            halfx8 f1[2], f2[2];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                int      r  = 8 * h + t;
                int      m  = (k * r) & 15;
                float    c  = (float)cos32(2 * m);
                float    s  = -(float)sin32(2 * m);
                float    v1 = i < 16 ? c : s;
                float    v2 = i < 16 ? -s : c;
                _Float16 h1 = (_Float16)v1, h2 = (_Float16)v2;
                f1[0][t] = h1, f1[1][t] = (_Float16)(v1 - (float)h1);
                f2[0][t] = h2, f2[1][t] = (_Float16)(v2 - (float)h2);
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                a1[q] = __builtin_bit_cast(u32x4, f1[q]);
                a2[q] = __builtin_bit_cast(u32x4, f2[q]);
            }
        } else {
            bf16x8 f1[3], f2[3];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                int   r  = 8 * h + t;              // input index (column of F)
                int   m  = (k * r) & 15;           // exp(-2*pi*i*m/16)
                float c  = (float)cos32(2 * m);    // cos(2*pi*m/16)
                float s  = -(float)sin32(2 * m);   // -sin(2*pi*m/16)
                float v1 = i < 16 ? c : s;         // [Fr; Fi]
                float v2 = i < 16 ? -s : c;        // [-Fi; Fr]
                __bf16 p0, p1, p2;
                split3(v1, p0, p1, p2);
                f1[0][t] = p0, f1[1][t] = p1, f1[2][t] = p2;
                split3(v2, p0, p1, p2);
                f2[0][t] = p0, f2[1][t] = p1, f2[2][t] = p2;
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                a1[q] = __builtin_bit_cast(u32x4, f1[q]);
                a2[q] = __builtin_bit_cast(u32x4, f2[q]);
            }
        }
    }

    // 32 butterflies: xr/xi hold this lane's 8 inputs of its butterfly (rows 8h .. 8h+7).
    template<int PREC = PREC_BF16X3>
    __device__ floatx16 run(const float* xr, const float* xi) const {
        floatx16 d = {};
        if constexpr (PREC == PREC_FP16X2) {
            halfx8 rh, rl, ih, il;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                _Float16 hr = (_Float16)xr[t], hi = (_Float16)xi[t];
                rh[t] = hr, rl[t] = (_Float16)(xr[t] - (float)hr);
                ih[t] = hi, il[t] = (_Float16)(xi[t] - (float)hi);
            }
            const halfx8 A1h = __builtin_bit_cast(halfx8, a1[0]), A1l = __builtin_bit_cast(halfx8, a1[1]);
            const halfx8 A2h = __builtin_bit_cast(halfx8, a2[0]), A2l = __builtin_bit_cast(halfx8, a2[1]);
            // smallest terms first
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A1l, rh, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A2l, ih, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A1h, rl, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A2h, il, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A1h, rh, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A2h, ih, d, 0, 0, 0);
            return d;
        } else {
            u32x4 wr[3], wi[3];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t x0, x1, x2, y0, y1, y2;
                split3_pk(xr[2 * t], xr[2 * t + 1], x0, x1, x2);
                split3_pk(xi[2 * t], xi[2 * t + 1], y0, y1, y2);
                wr[0][t] = x0, wr[1][t] = x1, wr[2][t] = x2;
                wi[0][t] = y0, wi[1][t] = y1, wi[2][t] = y2;
            }
            bf16x8 br[3], bi[3], A1[3], A2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                br[k] = __builtin_bit_cast(bf16x8, wr[k]);
                bi[k] = __builtin_bit_cast(bf16x8, wi[k]);
                A1[k] = __builtin_bit_cast(bf16x8, a1[k]);
                A2[k] = __builtin_bit_cast(bf16x8, a2[k]);
            }
            // smallest terms first
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[1], br[1], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[1], bi[1], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[0], br[2], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[0], bi[2], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[2], br[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[2], bi[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[0], br[1], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[0], bi[1], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[1], br[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[1], bi[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[0], br[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[0], bi[0], d, 0, 0, 0);
            return d;
        }
    }

    // In-place DFT16 of v[0..15] for every lane (all 64 lanes of the wavefront take part).
    // v_permlane32_swap(a, b) returns ([a_lo | b_lo], [a_hi | b_hi]) (lanes 0-31 | 32-63), so one
    // swap per value pair moves the butterflies into the MFMA layout and back, with no selects.
    template<int PREC = PREC_BF16X3>
    __device__ void dft16(cf* v) const {
        // group 0 = butterflies of lanes 0..31, group 1 = lanes 32..63; low lanes supply rows
        // 0..7 of a group's butterfly, high lanes rows 8..15
        float g0r[8], g0i[8], g1r[8], g1i[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            auto sr = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t].x), __float_as_uint(v[8 + t].x), false,
                                                       false);
            auto si = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t].y), __float_as_uint(v[8 + t].y), false,
                                                       false);
            g0r[t] = __uint_as_float(sr[0]), g1r[t] = __uint_as_float(sr[1]);
            g0i[t] = __uint_as_float(si[0]), g1i[t] = __uint_as_float(si[1]);
        }
        float u0 = 1.f, u1 = 1.f; // output scales 2^-e of the two butterflies in this lane's column
        if constexpr (PREC == PREC_FP16X2) {
            float s0, s1;
            col_scale(g0r, g0i, s0, u0);
            col_scale(g1r, g1i, s1, u1);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                g0r[t] *= s0, g0i[t] *= s0;
                g1r[t] *= s1, g1i[t] *= s1;
            }
        }
        floatx16 d0 = run<PREC>(g0r, g0i);
        floatx16 d1 = run<PREC>(g1r, g1i);
        // accumulator element q < 8 of lane l holds rows k + 4h (re; q + 8: im), k = 8*(q/4) + q%4,
        // of column l%32: after the swap every lane has rows k and k + 4 of its own butterfly
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int   k  = 8 * (q >> 2) + (q & 3);
            float r0 = d0[q], r1 = d1[q], i0 = d0[q + 8], i1 = d1[q + 8];
            if constexpr (PREC == PREC_FP16X2)
                r0 *= u0, r1 *= u1, i0 *= u0, i1 *= u1;
            auto sr = __builtin_amdgcn_permlane32_swap(__float_as_uint(r0), __float_as_uint(r1), false, false);
            auto si = __builtin_amdgcn_permlane32_swap(__float_as_uint(i0), __float_as_uint(i1), false, false);
            v[k]     = cf{__uint_as_float(sr[0]), __uint_as_float(si[0])};
            v[k + 4] = cf{__uint_as_float(sr[1]), __uint_as_float(si[1])};
        }
    }
};

} // namespace pcxk
