// fft_kernels.hpp -- gfx950 kernels of the FP32 FFT path.
//
//  k_row   FFTs along contiguous rows: the single-pass transform for N <= 16384 (one row per
//          transform), pass B of the forward four-step and pass 1 of the inverse four-step.
//  k_col   FFTs down columns of a row-major tile (stride N1): pass A of the forward four-step and
//          pass 2 of the inverse.  A wavefront reads and writes C consecutive floats per row.
//  k_twiddles, k_generate  plan-time twiddle tables (FP64, rounded to FP32) and seeded inputs.
//
// Inverse transforms run as conj(FFT(conj(x))): conj_in / conj_out negate imaginary parts at
// load and store, so every kernel only implements the forward butterflies.
#pragma once
#include "stockham.hpp"

namespace pcxk {

enum : int { LD_CONTIG = 0, LD_PERM = 1, LD_STRIDED = 2 };
enum : int { ST_CONTIG = 0, ST_PERM = 1, ST_STRIDED = 2 };

struct RowArgs {
    const float*  src;
    float*        dst;
    int64_t       src_dist, dst_dist; // floats between consecutive transforms
    int64_t       src_size;           // elements >= src_size of a transform read as zero
    int64_t       nrows;              // rows in the launch (transforms << rows_lg)
    int           src_lp, dst_lp;     // log2 of the pack size
    int           rows_lg;            // log2(rows per transform): 0 single pass, log2(N2) four-step
    int           load_kind, store_kind, order;
    int           conj_in, conj_out;
    float         scale;              // applied at the store
    const float2* twA;                // W_A, A = row length
    const float2* twl;                // optional post-FFT twiddle W_N^{a*row} = twl[e & lmask] * twh[e >> ls]
    const float2* twh;
    int           tw_ls;
};

struct ColArgs {
    const float*  src;
    float*        dst;
    int64_t       src_dist, dst_dist;
    int64_t       src_size;
    int64_t       ntiles; // transforms * (N1 / C)
    int           src_lp, dst_lp;
    int           n1_lg;  // log2 of the row length N1 (the column stride)
    int           conj_in, conj_out;
    float         scale;
    const float2* twA;    // W_A, A = column length N2
    const float2* twl;    // optional post-FFT twiddle W_N^{n1*a}
    const float2* twh;
    int           tw_ls;
};

__device__ __forceinline__ cf ld_cx(const float* __restrict__ base, int64_t i, int lp) {
    int64_t o = pidx(i, lp);
    return {base[o], base[o + ((int64_t)1 << lp)]};
}
__device__ __forceinline__ void st_cx(float* __restrict__ base, int64_t i, int lp, cf v) {
    int64_t o                       = pidx(i, lp);
    base[o]                         = v.x;
    base[o + ((int64_t)1 << lp)]    = v.y;
}
__device__ __forceinline__ cf tw2(const float2* __restrict__ twl, const float2* __restrict__ twh, int ls,
                                  uint32_t e) {
    return cmul(tw_load(twl, e & ((1u << ls) - 1)), tw_load(twh, e >> ls));
}

// Default rows per workgroup: enough rows for 256 threads.
template<int LA>
constexpr int default_rows() {
    return Plan<LA>::TA >= 256 ? 1 : 256 / Plan<LA>::TA;
}

template<int LA, int G_>
struct RowShape {
    static constexpr int TA = Plan<LA>::TA;
    static constexpr int G  = G_; // rows per workgroup
    static constexpr int NT = G * TA;
    // odd row stride: rows of a strided access land in distinct banks
    static constexpr int STRIDE = (1 << LA) + ((1 << LA) >> 5) + 1;
};

template<int LA, int G_>
__global__ __launch_bounds__(G_ * Plan<LA>::TA) void k_row(RowArgs p) {
    using P             = Plan<LA>;
    using S             = RowShape<LA, G_>;
    constexpr int A     = P::A;
    constexpr int TA    = P::TA;
    constexpr int PPT   = P::PPT;
    constexpr int G     = S::G;
    constexpr int NT    = S::NT;
    constexpr int STR   = S::STRIDE;
    constexpr int R0    = P::first_radix;
    constexpr int RL    = P::last_radix;
    __shared__ float lre[G * STR];
    __shared__ float lim[G * STR];

    const int     t       = threadIdx.x;
    const int     g       = t / TA;
    const int     jt      = t % TA;
    const int64_t row0    = (int64_t)blockIdx.x * G;
    const int64_t row     = row0 + g;
    const bool    row_ok  = row < p.nrows;
    const int64_t rmask   = ((int64_t)1 << p.rows_lg) - 1;
    const int64_t bt      = row >> p.rows_lg;
    const int64_t k2      = row & rmask;
    auto          addr    = [&](int a) { return g * STR + a + (a >> 5); };
    const float*  sbase   = p.src + bt * p.src_dist;
    float*        dbase   = p.dst + bt * p.dst_dist;

    cf v[PPT];
    // ------------------------------------------------------------------ load
    if (p.load_kind == LD_CONTIG) {
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                int64_t i        = in_elem<LA>(jt, m, r) + (k2 << LA);
                v[m * R0 + r]    = (row_ok && i < p.src_size) ? ld_cx(sbase, i, p.src_lp) : cf{0.f, 0.f};
            }
    } else {
        if (p.load_kind == LD_PERM) {
            // this row's spectrum occupies slots [A*brev(k2), A*brev(k2) + A); local slot s holds
            // element slot_freq(order, s)
            int64_t base = (int64_t)brev_bits((uint32_t)k2, p.rows_lg) << LA;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int s     = jt + u * TA;
                int e     = (int)slot_freq(p.order, (uint32_t)s, LA);
                cf  x     = row_ok ? ld_cx(sbase, base + s, p.src_lp) : cf{0.f, 0.f};
                int o     = addr(e);
                lre[o]    = x.x;
                lim[o]    = x.y;
            }
        } else { // LD_STRIDED: element k1 of row k2 sits at slot k2 + (k1 << rows_lg)
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int     idx  = t + u * NT;
                int     gg   = idx % G;
                int     k1   = idx / G;
                int64_t rw   = row0 + gg;
                cf      x    = {0.f, 0.f};
                if (rw < p.nrows) {
                    const float* b = p.src + (rw >> p.rows_lg) * p.src_dist;
                    x              = ld_cx(b, (rw & rmask) + ((int64_t)k1 << p.rows_lg), p.src_lp);
                }
                int o  = gg * STR + k1 + (k1 >> 5);
                lre[o] = x.x;
                lim[o] = x.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                int o         = addr(in_elem<LA>(jt, m, r));
                v[m * R0 + r] = cf{lre[o], lim[o]};
            }
    }
    if (p.conj_in) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    // ------------------------------------------------------------------ transform
    // (the barrier before the first stage's LDS writes also orders them after the reads above)
    stockham<LA, false>(v, lre, lim, jt, addr, p.twA);

    if (p.twl) {
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t e    = (uint32_t)out_elem<LA>(jt, m, r) * (uint32_t)k2;
                v[m * RL + r] = cmul(v[m * RL + r], tw2(p.twl, p.twh, p.tw_ls, e));
            }
    }
    {
        float sy = p.conj_out ? -p.scale : p.scale;
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            v[q].x *= p.scale;
            v[q].y *= sy;
        }
    }
    // ------------------------------------------------------------------ store
    if (p.store_kind == ST_CONTIG) {
        if (row_ok) {
#pragma unroll
            for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    int64_t i = out_elem<LA>(jt, m, r) + (k2 << LA);
                    st_cx(dbase, i, p.dst_lp, v[m * RL + r]);
                }
        }
        return;
    }
    __syncthreads(); // the last stage's LDS reads are done
#pragma unroll
    for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            int o  = addr(out_elem<LA>(jt, m, r));
            lre[o] = v[m * RL + r].x;
            lim[o] = v[m * RL + r].y;
        }
    __syncthreads();
    if (p.store_kind == ST_PERM) {
        if (row_ok) {
            int64_t base = (int64_t)brev_bits((uint32_t)k2, p.rows_lg) << LA;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int s = jt + u * TA;
                int o = addr((int)slot_freq(p.order, (uint32_t)s, LA));
                st_cx(dbase, base + s, p.dst_lp, cf{lre[o], lim[o]});
            }
        }
    } else { // ST_STRIDED: element k1 of row k2 goes to slot k2 + (k1 << rows_lg)
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            int     idx = t + u * NT;
            int     gg  = idx % G;
            int     k1  = idx / G;
            int64_t rw  = row0 + gg;
            if (rw < p.nrows) {
                int o    = gg * STR + k1 + (k1 >> 5);
                float* b = p.dst + (rw >> p.rows_lg) * p.dst_dist;
                st_cx(b, (rw & rmask) + ((int64_t)k1 << p.rows_lg), p.dst_lp, cf{lre[o], lim[o]});
            }
        }
    }
}

template<int LA, int C>
__global__ __launch_bounds__(C * Plan<LA>::TA) void k_col(ColArgs p) {
    using P           = Plan<LA>;
    constexpr int PPT = P::PPT;
    constexpr int R0  = P::first_radix;
    constexpr int RL  = P::last_radix;
    constexpr int A   = P::A;
    __shared__ float lre[A * C];
    __shared__ float lim[A * C];

    const int     t     = threadIdx.x;
    const int     c     = t % C;
    const int     jt    = t / C;
    const int     ctl   = p.n1_lg - ilog2c(C); // log2(column tiles per transform)
    const int64_t tile  = blockIdx.x;
    const int64_t bt    = tile >> ctl;
    const int64_t n1    = ((tile & (((int64_t)1 << ctl) - 1)) * C) + c;
    auto          addr  = [&](int a) { return a * C + c; };
    const float*  sbase = p.src + bt * p.src_dist;
    float*        dbase = p.dst + bt * p.dst_dist;

    cf v[PPT];
#pragma unroll
    for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            int64_t i     = n1 + ((int64_t)in_elem<LA>(jt, m, r) << p.n1_lg);
            v[m * R0 + r] = i < p.src_size ? ld_cx(sbase, i, p.src_lp) : cf{0.f, 0.f};
        }
    if (p.conj_in) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    stockham<LA, false>(v, lre, lim, jt, addr, p.twA);
    if (p.twl) {
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t e    = (uint32_t)n1 * (uint32_t)out_elem<LA>(jt, m, r);
                v[m * RL + r] = cmul(v[m * RL + r], tw2(p.twl, p.twh, p.tw_ls, e));
            }
    }
    float sy = p.conj_out ? -p.scale : p.scale;
#pragma unroll
    for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            cf      x = v[m * RL + r];
            int64_t i = n1 + ((int64_t)out_elem<LA>(jt, m, r) << p.n1_lg);
            st_cx(dbase, i, p.dst_lp, cf{x.x * p.scale, x.y * sy});
        }
}

// t[e] = W_M^(e*stride) = exp(-2*pi*i*e*stride/M) for e < count, computed in FP64, rounded to FP32.
__global__ void k_twiddles(float2* t, int64_t M, int64_t stride, int64_t count) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count)
        return;
    int64_t k = (e * stride) % M;
    double  s, c;
    sincospi(-2.0 * (double)k / (double)M, &s, &c);
    t[e] = make_float2((float)c, (float)s);
}

// Seeded inputs: output k of splitmix64(seed) is mix(seed + (k+1)*gamma), so element i uses
// outputs 2i and 2i+1: uniform [-1,1) re/im, or Box-Muller normal(0, sigma^2).
__device__ __forceinline__ uint64_t smix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__global__ void k_generate(float* dst, int lp, int64_t n, uint64_t seed, int dist, double sigma) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t gam = 0x9E3779B97F4A7C15ULL;
    double         u1  = (double)(smix(seed + (uint64_t)(2 * i + 1) * gam) >> 11) * (1.0 / 9007199254740992.0);
    double         u2  = (double)(smix(seed + (uint64_t)(2 * i + 2) * gam) >> 11) * (1.0 / 9007199254740992.0);
    float          re, im;
    if (dist == 0) {
        re = (float)(2.0 * u1 - 1.0);
        im = (float)(2.0 * u2 - 1.0);
    } else {
        double r = sqrt(-2.0 * log(1.0 - u1));
        double s, c;
        sincospi(2.0 * u2, &s, &c);
        re = (float)(sigma * r * c);
        im = (float)(sigma * r * s);
    }
    st_cx(dst, i, lp, cf{re, im});
}

} // namespace pcxk
