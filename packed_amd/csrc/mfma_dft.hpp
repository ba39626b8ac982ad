// mfma_dft.hpp -- radix-16 DFT on the gfx950 matrix cores with split-precision operands.
//
// A wavefront's 64 radix-16 butterflies (one per lane, 16 complex values in registers, the layout
// of the FP32 path) are computed as two real 32x32x16 products per operand pair:
//   [Yr; Yi] = [Fr; Fi] * Xr + [-Fi; Fr] * Xi,  F[k][r] = exp(-2*pi*i*k*r/16),
// with the 32 outputs of 32 butterflies in one 32x32 FP32 accumulator
// (v_mfma_f32_32x32x16_bf16).  Operands are split into three BF16 terms by truncation, which is
// exact (8 + 8 + 8 bits cover the FP32 significand): x = x0 + x1 + x2, F = F0 + F1 + F2, and the
// six products with combined weight >= 2^-16 are accumulated (smallest first) in FP32:
//   F1x1 + F0x2 + F2x0 + F0x1 + F1x0 + F0x0;  the dropped terms are below 2^-24 relative.
//
// Lane layout of v_mfma_f32_32x32x16_{bf16,f16}: A (32x16) lane l holds A[l%32][8*(l/32) + t],
// B (16x32) lane l holds B[8*(l/32) + t][l%32], t < 8; the FP32 accumulator lane l holds
// D[8*(v/4) + 4*(l/32) + v%4][l%32], v < 16.  Two half-wave exchanges (lane l <-> l^32) move the
// butterflies between the one-butterfly-per-lane layout and this one.
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
        bf16x8 br[3], bi[3];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            __bf16 p0, p1, p2;
            split3(xr[t], p0, p1, p2);
            br[0][t] = p0, br[1][t] = p1, br[2][t] = p2;
            split3(xi[t], p0, p1, p2);
            bi[0][t] = p0, bi[1][t] = p1, bi[2][t] = p2;
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
    __device__ void dft16(cf* v) const {
        const int h = (threadIdx.x >> 5) & 1;
        // group 0 = butterflies of lanes 0..31, group 1 = lanes 32..63; each lane supplies rows
        // 8h .. 8h+7 of one butterfly per group
        float g0r[8], g0i[8], g1r[8], g1i[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            cf    send = h ? v[t] : v[8 + t];
            float rr   = __shfl_xor(send.x, 32);
            float ri   = __shfl_xor(send.y, 32);
            if (h == 0) {
                g0r[t] = v[t].x, g0i[t] = v[t].y; // own butterfly, rows 0..7
                g1r[t] = rr, g1i[t] = ri;         // partner's butterfly, rows 0..7
            } else {
                g0r[t] = rr, g0i[t] = ri;                 // partner's butterfly, rows 8..15
                g1r[t] = v[8 + t].x, g1i[t] = v[8 + t].y; // own butterfly, rows 8..15
            }
        }
        floatx16 d0 = run(g0r, g0i);
        floatx16 d1 = run(g1r, g1i);
        // accumulator v: row 8*(v/4) + 4h + v%4 -> Yr[k] (v < 8) or Yi[k] (v >= 8) with
        // k = 8*((v/4)&1) + 4h + v%4; lane keeps its own butterfly's rows, swaps the rest
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int   k     = 8 * (q >> 2) + (q & 3); // + 4h
            float own_r = h ? d1[q] : d0[q];
            float own_i = h ? d1[q + 8] : d0[q + 8];
            float snd_r = h ? d0[q] : d1[q];
            float snd_i = h ? d0[q + 8] : d1[q + 8];
            float rr    = __shfl_xor(snd_r, 32);
            float ri    = __shfl_xor(snd_i, 32);
            // own rows: k + 4h; received rows: k + 4(1-h)
            if (h == 0) {
                v[k]     = cf{own_r, own_i};
                v[k + 4] = cf{rr, ri};
            } else {
                v[k + 4] = cf{own_r, own_i};
                v[k]     = cf{rr, ri};
            }
        }
    }
};

} // namespace pcxk
