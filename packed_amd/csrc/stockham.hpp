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
// this thread's transform in the split real/imaginary arrays `lre`/`lim`.  It is called as
// addr(base, off) with a compile-time offset, so LDS accesses use immediate offsets from one
// address per butterfly (the padded layouts are carry-free for every plan: pad(b + o) =
// pad(b) + pad(o) for the offsets a stage adds).
//
// Stage twiddles come from an LDS copy of W_A (TwTable): the whole table for A <= 1024, else a
// two-level split W_A^e = lo[e mod 64] * hi[e / 64] (one extra complex multiply, 2.5 KiB at
// A = 16384).  The table is filled once per workgroup from the FP32 table in global memory.
#pragma once
#include "dev_common.hpp"
#include "mfma_dft.hpp"

namespace pcxk {

// Plan options PX: a points-per-thread override in the low byte (the latency-bound fused
// kernels run 2^14 with 16), kHalfExchange to take the one-array exchange (Plan::HALF) below
// the size where it is the default, and kTwChain (below).
constexpr int kHalfExchange = 0x200;
// kTwChain: stage twiddle exponents by one addition per twiddle (TwTable::Seq) instead of one
// multiply each.  For kernels that run one transform per workgroup (no loop to hoist the
// per-thread twiddle addresses out of, and no registers to keep them in).
constexpr int kTwChain = 0x400;
template<int LA, bool REV = false, class T = float, int PX = 0>
struct Plan {
    static constexpr bool F64    = sizeof(T) == 8;
    static constexpr int A       = 1 << LA;
    static constexpr int PXV     = PX & 0xFF;
    static constexpr int NFULL   = LA / 4;           // radix-16 stages
    static constexpr int REM     = LA % 4;           // bits left for the small-radix stage
    static constexpr int NSTAGES = NFULL + (REM ? 1 : 0);
    // points per thread: more for the largest sizes, so that a 2^13 transform takes 256 threads
    // and a 2^14 transform 256 threads as well (two workgroups share a CU: one loads or stores
    // while the other computes).  FP64 values take two VGPRs each: at most 32 points per thread.
    static constexpr int PPT     = PXV > 0 ? PXV
                                 : F64  ? (LA >= 13 ? 32 : (LA >= 4 ? 16 : A))
                                        : (LA >= 14 ? 64 : LA >= 13 ? 32 : (LA >= 4 ? 16 : A));
    static constexpr int TA      = A / PPT;          // threads per transform
    // HALF: the stage exchange moves real parts, then imaginary parts, through one A-element LDS
    // array (2^14 FP32 points: 66 KiB instead of 132 KiB, so two workgroups fit a CU; FP64 from
    // 2^12 points, or wherever PX asks for it)
    static constexpr bool HALF   = (PX & kHalfExchange) || (F64 ? LA >= 12 : LA >= 14);
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

// LDS twiddle table of W_A^e = exp(-2*pi*i*e/A), elements of type T.
template<int LA, class T = float>
struct TwTable {
    static constexpr bool direct = LA <= (sizeof(T) == 8 ? 8 : 10);
    static constexpr int  LOB    = direct ? LA : 6;
    static constexpr int  NLO    = 1 << LOB;
    static constexpr int  NHI    = direct ? 0 : 1 << (LA - LOB);
    static constexpr int  FLOATS = 2 * (NLO + NHI); // elements: lo re, lo im, hi re, hi im

    // tw: global W_A table (interleaved re, im); nthreads threads cooperate.
    __device__ static void fill(T* t, const vec2<T>* __restrict__ tw, int tid, int nthreads) {
        for (int e = tid; e < NLO; e += nthreads) {
            vec2<T> w      = tw[e];
            t[e]           = w.x;
            t[NLO + e]     = w.y;
        }
        for (int h = tid; h < NHI; h += nthreads) {
            vec2<T> w               = tw[h << LOB];
            t[2 * NLO + h]          = w.x;
            t[2 * NLO + NHI + h]    = w.y;
        }
    }
    // The twiddles W^(r * e1), r = 1, 2, ..., of one butterfly, e1 = k << SH.  The exponent advances
    // by one addition per twiddle (kept opaque, so it is not folded back into a multiply per r) and
    // is held in bytes: one address per table lookup, with the imaginary parts and the high factor
    // at immediate offsets.  From SH = LOB on the low factor is W^0 = 1 (exactly): one lookup.
    template<int SH>
    struct Seq {
        static constexpr uint32_t S = sizeof(T);
        const char*               tb;
        uint32_t                  e, step; // exponent * S
        __device__ Seq(const T* t, uint32_t k) : tb((const char*)t), e(0), step((k << SH) * S) {}
        __device__ static T at(const char* p) { return *(const T*)p; }
        __device__ cx<T> next() {
            e += step;
            asm volatile("" : "+v"(e));
            if constexpr (direct) {
                return {at(tb + e), at(tb + NLO * S + e)};
            } else if constexpr (SH >= LOB) {
                const uint32_t h = e >> LOB;
                return {at(tb + 2 * NLO * S + h), at(tb + (2 * NLO + NHI) * S + h)};
            } else {
                const uint32_t l = e & ((NLO - 1) * S), h = (e >> LOB) & ~(S - 1);
                cx<T>          a = {at(tb + l), at(tb + NLO * S + l)};
                cx<T>          b = {at(tb + 2 * NLO * S + h), at(tb + (2 * NLO + NHI) * S + h)};
                return cmul(a, b);
            }
        }
    };
    // e a multiple of 2^SH: from SH = LOB on the low factor is W^0 = 1 (exactly), so one lookup
    template<int SH>
    __device__ static cx<T> get_sh(const T* t, uint32_t e) {
        if constexpr (!direct && SH >= LOB) {
            uint32_t hi = e >> LOB;
            return {t[2 * NLO + hi], t[2 * NLO + NHI + hi]};
        } else {
            return get(t, e);
        }
    }
    __device__ static cx<T> get(const T* t, uint32_t e) {
        if constexpr (direct) {
            return {t[e], t[NLO + e]};
        } else {
            uint32_t lo = e & (NLO - 1), hi = e >> LOB;
            cx<T>    a  = {t[lo], t[NLO + lo]};
            cx<T>    b  = {t[2 * NLO + hi], t[2 * NLO + NHI + hi]};
            return cmul(a, b);
        }
    }
};

// A copy of x that the compiler cannot see through: per-stage LDS addresses derived from it are
// computed where they are used instead of being hoisted (and spilled) for all stages at once.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}
// Scalar form: uniform values derived from it (offsets from a run-time pack size) are
// computed where used instead of all being held in SGPRs at once.
__device__ __forceinline__ int opaque_s(int x) {
    asm volatile("" : "+s"(x));
    return x;
}
// The same, also ordered after every memory access before it: addresses and predicates of a
// store phase are then computed after the last LDS exchange, not kept live across it.
__device__ __forceinline__ int opaque_after_memory(int x) {
    asm volatile("" : "+v"(x) : : "memory");
    return x;
}

// Reads of stage s (component C: 0 real, 1 imaginary, 2 both) from LDS into v.
template<int LA, bool REV, int s, int C, int PX = 0, class Addr, class T>
__device__ __forceinline__ void stage_read(cx<T>* v, const T* lre, const T* lim, int jt, const Addr& addr) {
    using P          = Plan<LA, REV, T, PX>;
    constexpr int R  = P::radix(s);
    constexpr int NB = P::PPT / R;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        int j = jt + m * P::TA;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int o = addr(j, r * (P::A / R));
            if constexpr (C != 1)
                v[m * R + r].x = lre[o];
            if constexpr (C != 0)
                v[m * R + r].y = lim[o];
        }
    }
}
// Writes of stage s (component C) from v to LDS.
template<int LA, bool REV, int s, int C, int PX = 0, class Addr, class T>
__device__ __forceinline__ void stage_write(const cx<T>* v, T* lre, T* lim, int jt, const Addr& addr) {
    using P          = Plan<LA, REV, T, PX>;
    constexpr int R  = P::radix(s);
    constexpr int Ns = P::ns(s);
    constexpr int NB = P::PPT / R;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        int j  = jt + m * P::TA;
        int hi = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int o = addr(hi, r * Ns);
            if constexpr (C != 1)
                lre[o] = v[m * R + r].x;
            if constexpr (C != 0)
                lim[o] = v[m * R + r].y;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// In-register exchange inside a wavefront.  Exchanging lane bit LB with a register-index bit
// (the value of lane l in register q and the value of lane l ^ 2^LB in register q ^ 2^i trade
// places when the two bits differ) moves each pair (a, b) = (v[q], v[q | 2^i]) of a lane with one
// v_permlane32_swap (LB = 5) or v_permlane16_swap (LB = 4) per dword.  Both sizes it serves,
// 2^9 (lane bit 4) and 2^10 (lane bits 4 and 5), then skip one of their two LDS exchanges.
// Lane bits 0..3 would need a partner fetch (DPP or ds_swizzle) and selects, three or four VALU
// instructions per dword and bit: on MI355X that lost to the LDS exchange (2^6: 1.18 against
// 0.95 ms, 2^7: 0.95 against 0.84 ms per 2 GiB batch; 2^5 unchanged), so those sizes keep LDS.
constexpr bool lane_exchange_on(int la) { return la == 9 || la == 10; }

template<int LB>
__device__ __forceinline__ void lane_swap(uint32_t& a, uint32_t& b) {
    static_assert(LB == 4 || LB == 5, "half-row / half-wave swaps only");
    if constexpr (LB == 5) {
        auto w = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a      = w[0];
        b      = w[1];
    } else {
        auto w = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a      = w[0];
        b      = w[1];
    }
}
template<int LB, class T>
__device__ __forceinline__ void lane_swap(T& a, T& b) {
    if constexpr (sizeof(T) == 4) {
        uint32_t x = __builtin_bit_cast(uint32_t, a), y = __builtin_bit_cast(uint32_t, b);
        lane_swap<LB>(x, y);
        a = __builtin_bit_cast(T, x), b = __builtin_bit_cast(T, y);
    } else {
        uint2 x = __builtin_bit_cast(uint2, a), y = __builtin_bit_cast(uint2, b);
        lane_swap<LB>(x.x, y.x);
        lane_swap<LB>(x.y, y.y);
        a = __builtin_bit_cast(T, x), b = __builtin_bit_cast(T, y);
    }
}
// Lane bits LB0 + i <-> register bits i (i < LR) of the PPT = 16 values of every lane.
template<int LB0, int LR, class T>
__device__ __forceinline__ void lane_exchange(cx<T>* v) {
    sfor<0, LR>([&](auto ii) {
        constexpr int i  = decltype(ii)::value;
        constexpr int LB = LB0 + i;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if ((q >> i & 1) == 0) {
                lane_swap<LB>(v[q].x, v[q | 1 << i].x);
                lane_swap<LB>(v[q].y, v[q | 1 << i].y);
            }
    });
}

// Matrix-core DFT16 of v (FP32 only; the FP64 path never selects it).
template<int PREC, class T>
__device__ __forceinline__ void dft16_mfma(const MfmaDft16* mf, cx<T>* v) {
    if constexpr (sizeof(T) == 4)
        mf->template dft16<PREC>(v);
}

// A thread's position in a stage: its butterfly index within the transform (jt) and the LDS
// address map of its transform (addr).
template<class Addr>
struct StageMap {
    int         jt;
    const Addr& addr;
};

// Twiddles W^(r e), r = 1 .. 15, of NB radix-16 butterflies at v, v + 16, ... (the same e) from a
// two-level table: W^e, W^2e, W^4e, W^8e by lookup (a product each), the other eleven as
// products of those (at most three deep), so four lookups instead of fifteen.  e < A / 16.
template<int LA, int NB, class T>
__device__ __forceinline__ void twiddle16_pow(cx<T>* v, const T* ltw, uint32_t e) {
    using TW       = TwTable<LA, T>;
    const cx<T> w1 = TW::get(ltw, e);
    const cx<T> w2 = TW::get(ltw, e << 1);
    const cx<T> w4 = TW::get(ltw, e << 2);
    const cx<T> w8 = TW::get(ltw, e << 3);
    auto        tw = [&](int r, cx<T> w) {
#pragma unroll
        for (int m = 0; m < NB; ++m)
            v[m * 16 + r] = cmul(v[m * 16 + r], w);
    };
    const cx<T> w6  = cmul(w4, w2);
    const cx<T> w10 = cmul(w8, w2);
    const cx<T> w12 = cmul(w8, w4);
    const cx<T> w14 = cmul(w12, w2);
    tw(1, w1);
    tw(2, w2);
    tw(3, cmul(w2, w1));
    tw(4, w4);
    tw(5, cmul(w4, w1));
    tw(6, w6);
    tw(7, cmul(w6, w1));
    tw(8, w8);
    tw(9, cmul(w8, w1));
    tw(10, w10);
    tw(11, cmul(w10, w1));
    tw(12, w12);
    tw(13, cmul(w12, w1));
    tw(14, w14);
    tw(15, cmul(w14, w1));
}

// PREC_BF16X3 / PREC_FP16X2 run the radix-16 butterflies on the matrix cores (mf holds the
// operands).
// With Plan::HALF, lre and lim are the same array and each exchange runs in two halves.
// T = double runs the same schedule in FP64 (PREC is then PREC_FP32, the plain butterflies).
// first / last (optional, not with Plan::HALF, at least two stages) give the first and the last
// stage another assignment of butterflies to threads than the middle stages (jt0, addr): the
// exchange before the last stage (or after the first) then also transposes, so a strided store
// (load) needs no LDS round trip of its own.
// LANES: the caller holds butterfly index jt0 in lane jt0 % 64 (jt0 = threadIdx.x % TA).  With 16
// points per thread and a small last radix RL, the exchange before the last stage is a swap of
// butterfly index bits [log2(TA/RL), log2 TA) with register bits [0, log2 RL).  With TA <= 64
// those index bits are lane bits [log2(TA/RL), log2 TA): for 2^9 and 2^10 the exchange runs in
// registers (lane_exchange), with no LDS round trip and no barrier.
// XLB >= 0 (with LANES, any TA): the caller placed butterfly index bits [log2(TA/RL), log2 TA)
// in lane bits [XLB, XLB + log2 RL) (4 or 5), so the same exchange runs in registers.
template<int LA, bool REV, class Addr, int PREC = PREC_FP32, class T = float, int PX = 0, bool FM = false,
         bool LM = false, class AddrF = Addr, class AddrL = Addr, bool LANES = false, int XLB = -1>
__device__ __forceinline__ void stockham(cx<T>* v, T* lre, T* lim, int jt0, const Addr& addr, const T* ltw,
                                         const MfmaDft16* mf = nullptr, const StageMap<AddrF>* first = nullptr,
                                         const StageMap<AddrL>* last = nullptr) {
    using P           = Plan<LA, REV, T, PX>;
    static_assert(PREC == PREC_FP32 || sizeof(T) == 4, "matrix-core butterflies are FP32 only");
    static_assert(!(FM || LM) || (!P::HALF && P::NSTAGES >= 2), "remapped stages need two stages and two arrays");
    constexpr int TA  = P::TA;
    constexpr int PPT = P::PPT;
    constexpr int RLX = P::radix(P::NSTAGES - 1);
    constexpr bool RX = LANES && !REV && !LM && !(FM && P::NSTAGES == 2) && PPT == 16 && P::NSTAGES >= 2 &&
                        RLX < 16 && (XLB >= 0 || (!P::HALF && TA <= 64 && lane_exchange_on(LA)));
    constexpr int  XL0 = XLB >= 0 ? XLB : ilog2c(TA) - ilog2c(RLX); // lane bit of index bit log2(TA/RL)
    sfor<0, P::NSTAGES>([&](auto ss) {
        constexpr int  s    = decltype(ss)::value;
        constexpr int  R    = P::radix(s);
        constexpr int  Ns   = P::ns(s);
        constexpr int  NB   = PPT / R; // butterflies per thread
        constexpr bool MAPF = FM && s == 0;
        constexpr bool MAPL = LM && s + 1 == P::NSTAGES;
        // (the 2^14-point plans need the registers; smaller plans run faster with the addresses
        // precomputed)
        int jt;
        if constexpr (P::HALF)
            jt = opaque(jt0);
        else if constexpr (MAPL)
            jt = last->jt;
        else if constexpr (MAPF)
            jt = first->jt;
        else
            jt = jt0;
        if constexpr (s > 0 && !P::HALF && !(RX && s + 1 == P::NSTAGES)) {
            __syncthreads(); // previous stage's LDS writes are visible
            if constexpr (MAPL)
                stage_read<LA, REV, s, 2, PX>(v, lre, lim, jt, last->addr);
            else
                stage_read<LA, REV, s, 2, PX>(v, lre, lim, jt, addr);
        }
        auto butterfly = [&](int m) {
            if constexpr (R == 16 && PREC != PREC_FP32)
                dft16_mfma<PREC>(mf, v + m * R);
            else
                dft<R>(v + m * R);
            if constexpr (NB > 1)
                __builtin_amdgcn_sched_barrier(0);
        };
        // Last stage with several butterflies per thread (j = jt + m*TA < Ns, Ns*R = A) under
        // kTwChain: W_A^(r*j) = W_A^(r*jt) * W_(A/TA)^(r*m), so the twiddles of m = 0 come from the
        // table and the others from them by compile-time rotations (twc).
        constexpr bool ROT = ((PX & kTwChain) != 0 || !TwTable<LA, T>::direct) && NB > 1 && Ns > 1 &&
                             Ns * R == P::A && PPT <= 32;
        // radix 16 on a two-level table below its high factor's granularity: twiddles by powers
        constexpr bool POW = R == 16 && Ns > 1 && !TwTable<LA, T>::direct &&
                             LA - ilog2c(Ns * R) < TwTable<LA, T>::LOB;
        if constexpr (ROT) {
            cx<T> bw[R];
            typename TwTable<LA, T>::template Seq<0> tw(ltw, (uint32_t)jt);
#pragma unroll
            for (int r = 1; r < R; ++r)
                bw[r] = tw.next();
            sfor<0, NB>([&](auto mm) {
                constexpr int m = decltype(mm)::value;
                sfor<1, R>([&](auto rr) {
                    constexpr int r = decltype(rr)::value;
                    v[m * R + r]    = cmul(v[m * R + r], twc<P::A / TA, r * m>(bw[r]));
                });
                butterfly(m);
            });
        } else if constexpr (POW && (NB == 1 || Ns <= TA)) {
            // every butterfly of the thread has the same k = jt mod Ns
            twiddle16_pow<LA, NB>(v, ltw, ((uint32_t)jt & (Ns - 1)) << (LA - ilog2c(Ns * R)));
#pragma unroll
            for (int m = 0; m < NB; ++m)
                butterfly(m);
        } else if constexpr ((PX & kTwChain) != 0 && NB > 1 && Ns > 1 && Ns <= TA) {
            // every butterfly of the thread has the same k = jt mod Ns: one twiddle sequence,
            // each twiddle applied to all of them while it is live
            constexpr int shift = LA - ilog2c(Ns * R);
            typename TwTable<LA, T>::template Seq<shift> tw(ltw, (uint32_t)jt & (Ns - 1));
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const cx<T> t = tw.next();
#pragma unroll
                for (int m = 0; m < NB; ++m)
                    v[m * R + r] = cmul(v[m * R + r], t);
            }
#pragma unroll
            for (int m = 0; m < NB; ++m)
                butterfly(m);
        } else {
            // twiddle and butterfly one radix-R group at a time, so only one group's twiddles are live
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                if constexpr (Ns > 1) {
                    constexpr int shift = LA - ilog2c(Ns * R);
                    uint32_t      k     = (uint32_t)(jt + m * TA) & (Ns - 1);
                    if constexpr (POW) {
                        twiddle16_pow<LA, 1>(v + m * R, ltw, k << shift);
                    } else if constexpr ((PX & kTwChain) != 0) {
                        typename TwTable<LA, T>::template Seq<shift> tw(ltw, k);
#pragma unroll
                        for (int r = 1; r < R; ++r)
                            v[m * R + r] = cmul(v[m * R + r], tw.next());
                    } else {
#pragma unroll
                        for (int r = 1; r < R; ++r)
                            v[m * R + r] = cmul(v[m * R + r],
                                                TwTable<LA, T>::template get_sh<shift>(ltw, (uint32_t)(r * k) << shift));
                    }
                }
                butterfly(m);
            }
        }
        if constexpr (RX && s + 2 == P::NSTAGES) {
            lane_exchange<XL0, ilog2c(RLX)>(v); // into the last stage's layout
        } else if constexpr (s + 1 < P::NSTAGES) {
            __syncthreads(); // every read of this stage is done before LDS is overwritten
            if constexpr (!P::HALF) {
                if constexpr (MAPF)
                    stage_write<LA, REV, s, 2, PX>(v, lre, lim, jt, first->addr);
                else
                    stage_write<LA, REV, s, 2, PX>(v, lre, lim, jt, addr);
            } else {
                // real parts out and back in the next stage's layout, then imaginary parts
                stage_write<LA, REV, s, 0, PX>(v, lre, lim, jt, addr);
                __syncthreads();
                stage_read<LA, REV, s + 1, 0, PX>(v, lre, lim, jt, addr);
                __syncthreads();
                stage_write<LA, REV, s, 1, PX>(v, lre, lim, jt, addr);
                __syncthreads();
                stage_read<LA, REV, s + 1, 1, PX>(v, lre, lim, jt, addr);
            }
        }
    });
}

// Element index held by register slot m*R + r on entry (first stage) and on exit (last stage).
template<int LA, bool REV = false, class T = float, int PX = 0>
__device__ __forceinline__ int in_elem(int jt, int m, int r) {
    using P = Plan<LA, REV, T, PX>;
    return jt + m * P::TA + r * (P::A / P::first_radix);
}
template<int LA, bool REV = false, class T = float, int PX = 0>
__device__ __forceinline__ int out_elem(int jt, int m, int r) {
    using P = Plan<LA, REV, T, PX>;
    return jt + m * P::TA + r * (P::A / P::last_radix);
}

} // namespace pcxk
