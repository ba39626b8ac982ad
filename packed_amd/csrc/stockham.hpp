// stockham.hpp -- LDS-staged Stockham autosort FFT core for gfx950.
//
// One transform of length A = 2^LA is held by TA = A / PPT threads, PPT points per thread.
// Stage s has radix R_s and Ns = R_0 * ... * R_{s-1}.  The radices are 16 plus one smaller
// radix for the remaining LA mod 4 bits, placed last (REV = false) or first (REV = true).
// Butterfly j in [0, A/R) reads a = j + r*(A/R), multiplies by W_{Ns*R}^{r*(j mod Ns)}, runs an
// in-register DFT_R and writes (j/Ns)*Ns*R + (j mod Ns) + r*Ns.  Thread jt owns butterflies
// j = jt + m*TA, m < PPT/R, so a wavefront always touches consecutive j.
//
// Entry: v[m*R0 + r] holds x[jt + m*TA + r*A/R0] (loaded by the caller, from HBM or LDS).
// Exit:  v[m*RL + r] holds X[jt + m*TA + r*A/RL] in natural order (RL = last radix).
// Between stages the data goes through LDS; `addr` maps an element index to a float offset of
// this thread's transform in the split real/imaginary arrays `lre`/`lim`.
#pragma once
#include "dev_common.hpp"

namespace pcxk {

template<int LA, bool REV = false>
struct Plan {
    static constexpr int A       = 1 << LA;
    static constexpr int NFULL   = LA / 4;           // radix-16 stages
    static constexpr int REM     = LA % 4;           // bits left for the small-radix stage
    static constexpr int NSTAGES = NFULL + (REM ? 1 : 0);
    static constexpr int PPT     = LA >= 4 ? 16 : A; // points per thread
    static constexpr int TA      = A / PPT;          // threads per transform
    static constexpr int radix(int s) {
        if (REM == 0)
            return 16;
        if (REV)
            return s == 0 ? (1 << REM) : 16;
        return s < NFULL ? 16 : (1 << REM);
    }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * radix(s - 1); }
    static constexpr int first_radix = radix(0);
    static constexpr int last_radix  = radix(NSTAGES - 1);
};

// Twiddle tables hold W_M^e = exp(-2*pi*i*e/M), e in [0, M), as interleaved (re, im) floats.
__device__ __forceinline__ cf tw_load(const float2* __restrict__ t, uint32_t e) {
    float2 w = t[e];
    return {w.x, w.y};
}

template<int LA, bool REV, class Addr>
__device__ __forceinline__ void stockham(cf* v, float* lre, float* lim, int jt, const Addr& addr,
                                         const float2* __restrict__ twA) {
    using P           = Plan<LA, REV>;
    constexpr int A   = P::A;
    constexpr int TA  = P::TA;
    constexpr int PPT = P::PPT;
    sfor<0, P::NSTAGES>([&](auto ss) {
        constexpr int s  = decltype(ss)::value;
        constexpr int R  = P::radix(s);
        constexpr int Ns = P::ns(s);
        constexpr int NB = PPT / R; // butterflies per thread
        if constexpr (s > 0) {
            __syncthreads(); // previous stage's LDS writes are visible
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                int j = jt + m * TA;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int o        = addr(j + r * (A / R));
                    v[m * R + r] = cf{lre[o], lim[o]};
                }
            }
        }
        if constexpr (Ns > 1) {
            constexpr int shift = LA - ilog2c(Ns * R);
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                uint32_t k = (uint32_t)(jt + m * TA) & (Ns - 1);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[m * R + r] = cmul(v[m * R + r], tw_load(twA, (uint32_t)(r * k) << shift));
            }
        }
#pragma unroll
        for (int m = 0; m < NB; ++m)
            dft<R>(v + m * R);
        if constexpr (s + 1 < P::NSTAGES) {
            __syncthreads(); // every read of this stage is done before LDS is overwritten
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                int j  = jt + m * TA;
                int hi = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int o  = addr(hi + r * Ns);
                    lre[o] = v[m * R + r].x;
                    lim[o] = v[m * R + r].y;
                }
            }
        }
    });
}

// Element index held by register slot m*R + r on entry (first stage) and on exit (last stage).
template<int LA, bool REV = false>
__device__ __forceinline__ int in_elem(int jt, int m, int r) {
    using P = Plan<LA, REV>;
    return jt + m * P::TA + r * (P::A / P::first_radix);
}
template<int LA, bool REV = false>
__device__ __forceinline__ int out_elem(int jt, int m, int r) {
    using P = Plan<LA, REV>;
    return jt + m * P::TA + r * (P::A / P::last_radix);
}

} // namespace pcxk
