// fxcorr_kernels.hpp -- fused overlap-save cross-correlation for gfx950.
//
// One workgroup row per block of the signal (fxcorr.hpp:102-136 runs these blocks one after
// another; here they are independent rows of one launch).  Per block:
//   load N samples (zero beyond the signal) -> forward Stockham FFT -> in registers,
//   Y = conj(X) * K with K = FFT(g)/N -> forward Stockham FFT of Y with the radix order
//   reversed -> conj -> store the first `step` outputs.
// conj(FFT(conj(X * conj(K)))) is the unnormalised inverse of X * conj(K), so the pointwise
// conj-multiply is fused between the last forward stage and the first inverse stage: the
// forward FFT ends with the small radix and the reversed plan starts with it, so both use the
// same register-to-element mapping and no LDS round trip sits between them.
#pragma once
#include "fft_kernels.hpp"

namespace pcxk {

// 64-bit packed element access (signals may exceed 2^31 floats).
__device__ __forceinline__ cf ld_cx64(const float* __restrict__ base, int64_t i, int lp) {
    int64_t o = pidx(i, lp);
    return {base[o], base[o + ((int64_t)1 << lp)]};
}
__device__ __forceinline__ void st_cx64(float* __restrict__ base, int64_t i, int lp, cf v) {
    int64_t o                    = pidx(i, lp);
    base[o]                      = v.x;
    base[o + ((int64_t)1 << lp)] = v.y;
}

struct XcArgs {
    const float* src;   // signal, pack src_lp
    float*       dst;   // output, pack dst_lp (may equal src: then halo holds the overlap)
    const float* hre;   // optional saved overlap: block b's samples [step, N) at b*hlen
    const float* him;
    const float* kre;   // kernel spectrum FFT(g)/N, natural order, split
    const float* kim;
    const float2* twA;  // W_N
    int64_t      len;   // signal length L
    int64_t      step;  // outputs kept per block
    int64_t      nblocks;
    int64_t      hlen;  // N - step
    int          src_lp, dst_lp;
    int64_t      src_dist, dst_dist; // floats between channels
    int64_t      channels;           // blocks of channel c are nblocks * c .. nblocks * (c + 1) - 1
    int64_t      blk0;               // first block of this launch (a launch covers < 2^32 work-items)
    int64_t      blk1;               // end of this launch's blocks
    int64_t      hblk0;              // block whose saved overlap is first in hre / him
    // In place without a saved overlap (N - step <= step): order[0] hands out workgroup tickets,
    // order[1 + w] is set once workgroup w has read its samples; a workgroup stores only after its
    // predecessor's flag (zeroed by k_xc_order_clear before each launch), or null.
    uint32_t*    order;
};

// Clears a launch's ticket counter and read flags (a kernel, so that graph replays clear too).
static __global__ __launch_bounds__(256) void k_xc_order_clear(uint32_t* o, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = 0u;
}

// Plan options of the correlation kernel.  The two transforms per block are bound by the
// latency of the LDS exchanges, and more waves per CU hide more of it:
//  * 2^14: 16 points per thread (1024 threads): L = 2^24, N = 2^14 runs in 0.135 ms against
//    0.138 ms with 32 points and 0.160 ms with 64; separate real / imaginary exchange arrays,
//    which leave one workgroup per CU, take 0.18 ms;
//  * 2^13: the one-array exchange (35 KiB, four workgroups per CU instead of two): 512 channels
//    of 2^20 samples with N = 2^13 take 2.65 ms against 3.52 ms;
//  * else the default rule;
//  * twiddle exponents by addition chains (kTwChain): one block per workgroup, so there is no
//    loop to hoist the per-thread twiddle addresses out of.
template<int LA>
constexpr int xc_px() {
    return kTwChain | (LA >= 14 ? 16 : (LA == 13 ? kHalfExchange : 0));
}
template<int LA>
constexpr int xc_threads() {
    return Plan<LA, false, float, xc_px<LA>()>::TA;
}
template<int LA>
constexpr int xc_rows() {
    return xc_threads<LA>() >= 256 ? 1 : 256 / xc_threads<LA>();
}
template<int LA, int G>
constexpr int xc_lds_floats() {
    return RowShape<LA, G, float, xc_px<LA>()>::DATA + TwTable<LA>::FLOATS;
}

// 2^14 blocks, 16 points per thread (xc_lanes; FP32 and matrix-core butterflies alike): the
// exchange before the radix-4 stage of both transforms runs in registers.  Thread t takes
// butterfly index xl_jt(t), t with bits 4, 5 and 8, 9 swapped, so index bits 8, 9 (the two that
// exchange trades with register bits 0, 1) sit in lane bits 4, 5, where v_permlane16_swap /
// v_permlane32_swap reach.  The other two exchanges stay in LDS, at element e's address
// xl_lds(e): e bits 0-3 -> 0-3, 8-9 -> 4-5, 10-11 -> 6-7, 4-7 -> 8-11, 12-13 -> 12-13, plus
// e bits 4-7 (as 0..15) and 16 * e bits 12-13.  The pad is linear in the bits (so base + offset
// addressing stays exact) and never decreases as the address bits above 5 count up (so no two
// elements share an address; the highest is 16446, inside the row stride); a wavefront's lanes
// vary element bits {0-3, 8, 9} in the stage reads, {4-7, 12, 13} in the stage-0 writes and
// {0-3, 12, 13} in the stage-1 writes, and each of these sets lands on 64 distinct banks.
template<int LA, int PPT, int RL, int PREC>
constexpr bool xc_lanes() {
    return LA == 14 && PPT == 16 && RL == 4;
}
__device__ __forceinline__ int xl_jt(int t) {
    return (t & ~0x330) | ((t >> 4) & 0x30) | ((t << 4) & 0x300);
}
__device__ __forceinline__ int xl_lds(int e) {
    return ((e & 0x300F) | ((e >> 4) & 0xF0) | ((e << 4) & 0xF00)) + ((e >> 4) & 0xF) + ((e >> 8) & 0x30);
}

// B32: the signal fits 32-bit buffer offsets (one SGPR descriptor, one VGPR per address).
// PREC_BF16X3 runs the radix-16 butterflies of both transforms on the matrix cores.
// ORD: in place in ticket order (XcArgs::order), a variant of its own so that the out-of-place
// kernel's register allocation does not change.
template<int LA, int G, bool B32, int PREC = PREC_FP32, bool ORD = false>
__global__ __launch_bounds__(G * xc_threads<LA>(),
                             PREC == PREC_FP32 ? waves_per_simd(G * xc_threads<LA>(), 4 * xc_lds_floats<LA, G>())
                                               : (waves_per_simd(G * xc_threads<LA>(), 4 * xc_lds_floats<LA, G>()) > 4
                                                      ? waves_per_simd(G * xc_threads<LA>(), 4 * xc_lds_floats<LA, G>()) / 2
                                                      : 2))
void k_fxcorr(XcArgs p) {
    constexpr int PX   = xc_px<LA>();
    using PF           = Plan<LA, false, float, PX>;
    constexpr bool XL  = xc_lanes<LA, PF::PPT, PF::last_radix, PREC>();
    // the inverse: the reversed plan (its first radix-RL stage takes the forward output's
    // register map), or with XL the forward plan again after a register renaming
    using PI           = Plan<LA, !XL, float, PX>;
    constexpr int TA   = PF::TA;
    constexpr int PPT  = PF::PPT;
    constexpr int R0   = PF::first_radix;
    constexpr int RL   = PF::last_radix;
    constexpr int RLI  = PI::last_radix;
    using RS           = RowShape<LA, G, float, PX>;
    constexpr int STR  = RS::STRIDE;
    static_assert(XL || PI::first_radix == RL, "fused multiply needs matching register maps");
    static_assert(!XL || (G == 1 && R0 == PPT), "register renaming assumes one block and one first-stage butterfly");
    __shared__ float lds[xc_lds_floats<LA, G>()];
    __shared__ uint32_t ticket;
    using LC             = LdsCarve<float, RS::DATA, TwTable<LA>::FLOATS, 0>;
    float*           ltw = lds + LC::TOFF;
    float*           lre = lds + LC::DOFF;
    float*           lim = RS::im(lre);
    if (ORD && threadIdx.x == 0)
        ticket = __hip_atomic_fetch_add(p.order, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    TwTable<LA>::fill(ltw, p.twA, threadIdx.x, G * TA);
    __syncthreads();
    // In place, workgroups take their blocks in ticket order (the order in which they started), so
    // the predecessor a workgroup waits for has started and reads its samples without waiting: the
    // wait always ends.  Tickets are spread over the XCDs as block indices are (xcd_item), except
    // that the first workgroup of each XCD's range then waits for the last of the previous range.
    const int64_t wgi = ORD ? xcd_map(__builtin_amdgcn_readfirstlane(ticket)) : xcd_item(0);

    MfmaDft16 mf;
    if constexpr (PREC != PREC_FP32)
        mf.template init<PREC>();
    const int     t    = threadIdx.x;
    // a wavefront lies in one block when TA >= 64: keep the block's offsets in SGPRs
    const int     g    = TA >= 64 ? __builtin_amdgcn_readfirstlane(t / TA) : t / TA;
    const int     jt   = XL ? xl_jt(t % TA) : t % TA;
    const int64_t bg   = p.blk0 + wgi * G + g; // block over all channels; neighbours share one L2
    const bool    ok   = bg < p.blk1;
    const int64_t ch   = ok ? bg / p.nblocks : 0;
    const int64_t b    = ok ? bg - ch * p.nblocks : 0;
    const int64_t off  = b * p.step;
    const float*  src  = p.src + ch * p.src_dist;
    float*        dst  = p.dst + ch * p.dst_dist;
    auto          addr = [&](int b, int o) {
        if constexpr (XL)
            return g * STR + xl_lds(b) + xl_lds(o);
        else
            return g * STR + b + (b >> 5) + o + (o >> 5);
    };

    // Fast addressing (B32, block offset a multiple of both packs): element off + a of the
    // signal sits at poff(off) + poff(a), so a resource based at poff(off) serves the block, and
    // its range check supplies the zeros past the signal (loads) and drops the outputs past
    // `keep` (stores): no per-element branches.
    const bool aligned = B32 && (off & (((int64_t)1 << p.src_lp) - 1)) == 0 &&
                         (off & (((int64_t)1 << p.dst_lp) - 1)) == 0;
    cf v[PPT];
    if (aligned && !p.hre) {
        const int64_t  rem = p.len - off;
        const uint32_t nv  = ok ? (uint32_t)(rem < Plan<LA>::A ? rem : Plan<LA>::A) : 0u;
        const BufR<>   sr(src + poff((uint32_t)off, p.src_lp), p.src_lp, poff(nv, p.src_lp));
        const uint32_t vo  = poff((uint32_t)jt, p.src_lp);
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                uint32_t c    = (uint32_t)(in_elem<LA, false, float, PX>(jt, m, r) - jt);
                v[m * R0 + r] = sr.ld(vo + poff(c, p.src_lp), 0);
            }
    } else {
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                int     a = in_elem<LA, false, float, PX>(jt, m, r);
                int64_t i = off + a;
                cf      x = {0.f, 0.f};
                if (ok && i < p.len) {
                    if (p.hre && a >= p.step) {
                        int64_t h = (bg - p.hblk0) * p.hlen + (a - p.step);
                        x         = {p.hre[h], p.him[h]};
                    } else if constexpr (B32) {
                        x = BufR<>(src, p.src_lp).ld(poff((uint32_t)i, p.src_lp), 0);
                    } else {
                        x = ld_cx64(src, i, p.src_lp);
                    }
                }
                v[m * R0 + r] = x;
            }
    }
    stockham<LA, false, decltype(addr), PREC, float, PX, false, false, decltype(addr), decltype(addr), true,
             XL ? 4 : -1>(v, lre, lim, jt, addr, ltw, &mf);
    // In place, the read flag goes up as early as here, except at 2^13, where the registers for
    // it are not there (it then goes up right before the wait).
    constexpr bool FLAG_EARLY = LA != 13;
    if constexpr (ORD && FLAG_EARLY) {
        // Every load of this workgroup has returned: each wave's values went through the LDS
        // exchange of the first stage, behind a barrier (one-stage plans: the barrier here), so
        // the successor may overwrite them.  No data passes through the flags, so they are relaxed
        // device-coherent accesses: an acquire or release at device scope would write back or
        // invalidate the XCD's whole L2.
        if constexpr (Plan<LA, false, float, PX>::NSTAGES == 1)
            __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_store(p.order + 1 + wgi, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    {
        // Y = conj(X) * K.  The spectrum is read in two halves: the second half's offset depends
        // (opaquely) on the first half's products, so at most half of it is in registers beside v.
        const BufR<> kr(p.kre, LA); // split spectrum: re[N], im[N]
        constexpr int NG = PPT / RL;
        int           jk = jt;
#pragma unroll
        for (int m = 0; m < NG; ++m) {
            if constexpr (NG >= 2 && PPT >= 32) {
                if (m == NG / 2)
                    asm volatile("" : "+v"(jk) : "v"(v[0].x), "v"(v[(NG / 2) * RL - 1].y));
            }
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t c    = (uint32_t)(out_elem<LA, false, float, PX>(jt, m, r) - jt);
                cf       k    = kr.ld((uint32_t)jk, c, 0);
                v[m * RL + r] = cmul(cconj(v[m * RL + r]), k);
            }
        }
    }
    if constexpr (XL) {
        // forward output slot m*RL + r holds element jt + (m + r*PPT/RL)*TA; the forward plan's
        // first stage takes element jt + k*TA in slot k
        constexpr int NG = PPT / RL;
        cf            u[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            u[k] = v[(k % NG) * RL + k / NG];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            v[k] = u[k];
        stockham<LA, false, decltype(addr), PREC, float, PX, false, false, decltype(addr), decltype(addr), true, 4>(
            v, lre, lim, jt, addr, ltw, &mf);
    } else {
        stockham<LA, true, decltype(addr), PREC, float, PX>(v, lre, lim, jt, addr, ltw, &mf);
    }
    const int64_t keep = p.len - off < p.step ? p.len - off : p.step;
    if constexpr (ORD) {
        // the first N - step samples of these blocks, which the predecessor reads, are
        // overwritten only after its flag (bounded wait: 2^22 polls, seconds, far beyond a
        // launch's run time); the flag is set without waiting, so every wait ends
        if constexpr (!FLAG_EARLY) {
            __builtin_amdgcn_s_waitcnt(0); // this wave's loads have returned; the barrier: every wave's
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if constexpr (!FLAG_EARLY)
                __hip_atomic_store(p.order + 1 + wgi, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (wgi > 0)
                for (uint32_t spin = 0; spin < (1u << 22); ++spin) {
                    if (__hip_atomic_load(p.order + wgi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                        break;
                    __builtin_amdgcn_s_sleep(2);
                }
        }
        __syncthreads();
    }
    if (aligned) {
        const int      jt = opaque_after_memory(XL ? xl_jt(t % TA) : t % TA);
        const BufW<>   dw(dst + poff((uint32_t)off, p.dst_lp), p.dst_lp, ok ? poff((uint32_t)keep, p.dst_lp) : 0u);
        const uint32_t vo = poff((uint32_t)jt, p.dst_lp);
#pragma unroll
        for (int m = 0; m < PPT / RLI; ++m)
#pragma unroll
            for (int r = 0; r < RLI; ++r) {
                uint32_t c = (uint32_t)(out_elem<LA, !XL, float, PX>(jt, m, r) - jt);
                dw.st(vo + poff(c, p.dst_lp), cconj(v[m * RLI + r]), 0);
            }
    } else if (ok) {
#pragma unroll
        for (int m = 0; m < PPT / RLI; ++m)
#pragma unroll
            for (int r = 0; r < RLI; ++r) {
                int a = out_elem<LA, !XL, float, PX>(jt, m, r);
                if (a < keep) {
                    if constexpr (B32)
                        BufW<>(dst, p.dst_lp).st(poff((uint32_t)(off + a), p.dst_lp), cconj(v[m * RLI + r]), 0);
                    else
                        st_cx64(dst, off + a, p.dst_lp, cconj(v[m * RLI + r]));
                }
            }
    }
}

// ---- block path for FFT sizes above 16384: the blocks go through HBM in a buffer of nb blocks
// (block j: re[N] then im[N] at 2*N*j), transformed by the batched four-step FFT.
struct XcBlockArgs {
    const float* src;  // signal, pack src_lp
    float*       dst;  // output, pack dst_lp
    float*       buf;  // nb blocks of 2N floats
    const float* kre;  // kernel spectrum FFT(g)/N, natural order, split
    const float* kim;
    int64_t      len, step, nblocks;
    int64_t      src_dist, dst_dist; // floats between channels
    int64_t      bg0, nb;            // global blocks bg0 .. bg0 + nb - 1 (channel-major)
    int          src_lp, dst_lp, lg;
    float        scale;              // 1/N (the kernel spectrum is FFT(g))
};

// buf[j] = the N samples from block bg0 + j's offset, zero past the signal.
static __global__ __launch_bounds__(256) void k_xc_gather(XcBlockArgs a) {
    const int64_t n     = (int64_t)1 << a.lg;
    const int64_t total = a.nb << a.lg;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx >> a.lg, i = idx & (n - 1);
        const int64_t bg = a.bg0 + j;
        const int64_t ch = bg / a.nblocks, b = bg - ch * a.nblocks;
        const int64_t s  = b * a.step + i;
        const cf      x  = s < a.len ? ld_cx64(a.src + ch * a.src_dist, s, a.src_lp) : cf{0.f, 0.f};
        a.buf[2 * n * j + i]     = x.x;
        a.buf[2 * n * j + n + i] = x.y;
    }
}
// buf[j] = buf[j] * conj(K) / N: the spectrum of the correlation block (fxcorr.hpp:67-68, 115).
static __global__ __launch_bounds__(256) void k_xc_mul(XcBlockArgs a) {
    const int64_t n     = (int64_t)1 << a.lg;
    const int64_t total = a.nb << a.lg;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx >> a.lg, i = idx & (n - 1);
        float*        re = a.buf + 2 * n * j + i;
        const cf      x  = {re[0], re[n]};
        const cf      y  = cmulc(x, cf{a.kre[i], a.kim[i]});
        re[0]            = y.x * a.scale;
        re[n]            = y.y * a.scale;
    }
}
// The first `step` outputs of block bg0 + j (fewer at the signal's end) to the destination.
static __global__ __launch_bounds__(256) void k_xc_scatter(XcBlockArgs a) {
    const int64_t n     = (int64_t)1 << a.lg;
    const int64_t total = a.nb * a.step;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / a.step, i = idx - j * a.step;
        const int64_t bg = a.bg0 + j;
        const int64_t ch = bg / a.nblocks, b = bg - ch * a.nblocks;
        const int64_t s  = b * a.step + i;
        if (s < a.len)
            st_cx64(a.dst + ch * a.dst_dist, s, a.dst_lp, cf{a.buf[2 * n * j + i], a.buf[2 * n * j + n + i]});
    }
}

// Saves, for the blocks bg0 .. bg0 + count - 1 (over all channels), the samples
// [b*step + step, b*step + N) that later blocks overwrite when the correlation runs in place.
static __global__ void k_fxcorr_halo(const float* src, int src_lp, int64_t src_dist, int64_t len, int64_t step, int64_t hlen,
                              int64_t nblocks, int64_t bg0, int64_t count, float* hre, float* him) {
    const int64_t total = count * hlen;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t bg = bg0 + idx / hlen, h = idx % hlen;
        int64_t ch = bg / nblocks, b = bg % nblocks;
        int64_t i  = b * step + step + h;
        cf      x  = i < len ? ld_cx64(src + ch * src_dist, i, src_lp) : cf{0.f, 0.f};
        hre[idx]   = x.x;
        him[idx]   = x.y;
    }
}

} // namespace pcxk
