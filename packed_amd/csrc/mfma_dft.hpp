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

enum : int { PREC_FP32 = 0, PREC_BF16X3 = 1 };

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

// Per-lane A operands: [Fr; Fi] and [-Fi; Fr], three split terms each.
struct MfmaDft16 {
    bf16x8 a1[3], a2[3];

    __device__ void init() {
        const int lane = threadIdx.x & 63;
        const int i    = lane & 31;
        const int h    = lane >> 5;
        const int k    = i & 15; // output frequency of this row
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
            a1[0][t] = p0, a1[1][t] = p1, a1[2][t] = p2;
            split3(v2, p0, p1, p2);
            a2[0][t] = p0, a2[1][t] = p1, a2[2][t] = p2;
        }
    }

    // 32 butterflies: xr/xi hold this lane's 8 inputs of its butterfly (rows 8h .. 8h+7).
    __device__ floatx16 run(const float* xr, const float* xi) const {
        u32x4 wr[3], wi[3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t a0, a1, a2, b0, b1, b2;
            split3_pk(xr[2 * t], xr[2 * t + 1], a0, a1, a2);
            split3_pk(xi[2 * t], xi[2 * t + 1], b0, b1, b2);
            wr[0][t] = a0, wr[1][t] = a1, wr[2][t] = a2;
            wi[0][t] = b0, wi[1][t] = b1, wi[2][t] = b2;
        }
        bf16x8 br[3], bi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            br[k] = __builtin_bit_cast(bf16x8, wr[k]);
            bi[k] = __builtin_bit_cast(bf16x8, wi[k]);
        }
        floatx16 d = {};
        // smallest terms first
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[1], br[1], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[1], bi[1], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[0], br[2], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[0], bi[2], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[2], br[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[2], bi[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[0], br[1], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[0], bi[1], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[1], br[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[1], bi[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[0], br[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[0], bi[0], d, 0, 0, 0);
        return d;
    }

    // In-place DFT16 of v[0..15] for every lane (all 64 lanes of the wavefront take part).
    // v_permlane32_swap(a, b) returns ([a_lo | b_lo], [a_hi | b_hi]) (lanes 0-31 | 32-63), so one
    // swap per value pair moves the butterflies into the MFMA layout and back, with no selects.
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
        floatx16 d0 = run(g0r, g0i);
        floatx16 d1 = run(g1r, g1i);
        // accumulator element q < 8 of lane l holds rows k + 4h (re; q + 8: im), k = 8*(q/4) + q%4,
        // of column l%32: after the swap every lane has rows k and k + 4 of its own butterfly
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int  k  = 8 * (q >> 2) + (q & 3);
            auto sr = __builtin_amdgcn_permlane32_swap(__float_as_uint(d0[q]), __float_as_uint(d1[q]), false, false);
            auto si =
                __builtin_amdgcn_permlane32_swap(__float_as_uint(d0[q + 8]), __float_as_uint(d1[q + 8]), false, false);
            v[k]     = cf{__uint_as_float(sr[0]), __uint_as_float(si[0])};
            v[k + 4] = cf{__uint_as_float(sr[1]), __uint_as_float(si[1])};
        }
    }
};

} // namespace pcxk
