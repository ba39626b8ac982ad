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
//
// Global addressing: one 64-bit base pointer per row (real part) and one for the imaginary part
// (base + P), then 32-bit packed offsets i + (i / P) * P (element_access.hpp:8-11).
#pragma once
#include "stockham.hpp"

namespace pcxk {

enum : int { LD_CONTIG = 0, LD_PERM = 1, LD_STRIDED = 2 };
enum : int { ST_CONTIG = 0, ST_PERM = 1, ST_STRIDED = 2 };

template<class T>
struct RowArgsT {
    const T*      src;
    T*            dst;
    int64_t       src_dist, dst_dist; // floats between consecutive transforms
    int64_t       src_size;           // elements >= src_size of a transform read as zero
    int64_t       nrows;              // rows in the launch (transforms << rows_lg)
    int           src_lp, dst_lp;     // log2 of the pack size
    int           rows_lg;            // log2(rows per transform): 0 single pass, log2(N2) four-step
    int           order;
    int           conj_in, conj_out;
    T             scale;              // applied at the store
    const vec2<T>* twA;               // W_A, A = row length
    const vec2<T>* twl;               // post-FFT twiddle W_N^{a*row} = twl[e & lmask] * twh[e >> ls]
    const vec2<T>* twh;
    int           tw_ls, tw_lg;       // ls and log2(N)
    int           nt_load, nt_store;  // non-temporal (streaming) global loads / stores
    const int64_t* src_offs;          // optional per-transform element offsets (instead of b * dist)
    const int64_t* dst_offs;
    // rsplit_lg > 0 (contiguous loads only): row r of a transform is read from row
    // ((r mod 2^rsplit_lg) << (rows_lg - rsplit_lg)) + (r >> rsplit_lg), i.e. the two parts of the
    // row index trade places (the last pass of the three-pass transform)
    int           rsplit_lg;
};
using RowArgs = RowArgsT<float>;

template<class T>
struct ColArgsT {
    const T*      src;
    T*            dst;
    int64_t       src_dist, dst_dist;
    int64_t       src_size;
    int64_t       ntiles; // transforms * (N1 / C)
    int           src_lp, dst_lp;
    int           n1_lg;  // log2 of the row length N1 (the column stride)
    int           conj_in, conj_out;
    T             scale;
    const vec2<T>* twA;   // W_A, A = column length N2
    const vec2<T>* twl;   // post-FFT twiddle W_N^{n1*a}
    const vec2<T>* twh;
    int           tw_ls, tw_lg;
    int           nt_load, nt_store;
    const int64_t* src_offs;
    const int64_t* dst_offs;
};
using ColArgs = ColArgsT<float>;

// XCD-aware order of work items.  Workgroup w runs on XCD w mod 8 (round-robin placement); items
// are dealt in rounds of gridDim.x, and within a round XCD x takes the contiguous range
// [x * gridDim.x / 8, (x + 1) * gridDim.x / 8), so neighbouring items -- which share 128-byte
// lines when a tile covers 64 bytes of a row, or input samples in overlap-save -- meet in one L2
// instead of being fetched from HBM by two XCDs.  (Placement only affects speed.)  Grids that are
// not a multiple of 8 keep the plain order.
__device__ __forceinline__ int64_t xcd_map(uint32_t i) {
    const uint32_t g = gridDim.x;
    if (g & 7u)
        return i;
    return (int64_t)(i & 7u) * (g >> 3) + (i >> 3);
}
__device__ __forceinline__ int64_t xcd_item(int64_t round_base) { return round_base + xcd_map(blockIdx.x); }

// Base of transform b: b * dist elements, or offs[b] when a per-transform offset table is given.
template<class T>
__device__ __forceinline__ const T* tbase(const T* base, const int64_t* offs, int64_t dist, int64_t b) {
    return offs ? base + offs[b] : base + b * dist;
}
template<class T>
__device__ __forceinline__ T* tbase(T* base, const int64_t* offs, int64_t dist, int64_t b) {
    return offs ? base + offs[b] : base + b * dist;
}

// 32-bit packed offset; lp is clamped so the shift stays defined (offsets are < 2^31).
__device__ __forceinline__ uint32_t poff(uint32_t i, int lp) {
    int l = lp < 30 ? lp : 30;
    return i + ((i >> l) << l);
}
// poff(jt + c) = poff(jt) + poff(c) when c is a multiple of the thread count TA of a transform
// and jt < TA (both powers of two, any power-of-two pack): the contiguous loads and stores of a
// thread share one per-lane offset poff(jt) and add the uniform poff(c) as a scalar offset.
template<class T>
__device__ __forceinline__ T ldf(const T* p, int nt) { return nt ? __builtin_nontemporal_load(p) : *p; }
template<class T>
__device__ __forceinline__ void stf(T* p, T v, int nt) {
    if (nt)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Re/im views of one packed vector.
template<class T = float>
struct CxPtr {
    const T* re;
    const T* im;
    __device__ CxPtr(const T* base, int lp) : re(base), im(base + ((int64_t)1 << lp)) {}
    __device__ cx<T> ld(uint32_t o, int nt) const { return {ldf(re + o, nt), ldf(im + o, nt)}; }
    __device__ cx<T> ld(uint32_t vo, uint32_t so, int nt) const { return ld(vo + so, nt); }
};
template<class T = float>
struct CxPtrW {
    T* re;
    T* im;
    __device__ CxPtrW(T* base, int lp) : re(base), im(base + ((int64_t)1 << lp)) {}
    __device__ void st(uint32_t o, cx<T> v, int nt) const {
        stf(re + o, v.x, nt);
        stf(im + o, v.y, nt);
    }
    __device__ void st(uint32_t vo, uint32_t so, cx<T> v, int nt) const { st(vo + so, v, nt); }
};

// Buffer-resource views (wave-uniform base in SGPRs, 32-bit byte offsets): one VGPR per address.
template<class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const T* base, uint32_t bytes = 0x7FFFFFFFu) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), 0, (int)bytes, 0x00020000);
}
// One element load / store through a buffer resource (byte offsets vb + sb).
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float buf_ld(float*, __amdgpu_buffer_rsrc_t r, uint32_t vb, uint32_t sb) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vb, sb, 0));
}
__device__ __forceinline__ double buf_ld(double*, __amdgpu_buffer_rsrc_t r, uint32_t vb, uint32_t sb) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vb, sb, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t vb, uint32_t sb) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vb, sb, 0);
}
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, uint32_t vb, uint32_t sb) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vb, sb, 0);
}
template<class T = float>
struct BufR {
    static constexpr uint32_t E = sizeof(T);
    __amdgpu_buffer_rsrc_t re, im;
    __device__ BufR(const T* base, int lp) : re(rsrc(base)), im(rsrc(base + ((int64_t)1 << lp))) {}
    // Loads at element offsets >= nfl return 0 (the buffer range check).  With nfl = poff(lim),
    // exactly the elements i >= lim read as zero, in both views (poff is increasing).
    __device__ BufR(const T* base, int lp, uint32_t nfl)
        : re(rsrc(base, nfl * E)), im(rsrc(base + ((int64_t)1 << lp), nfl * E)) {}
    __device__ cx<T> ld(uint32_t o, int) const {
        return {buf_ld((T*)nullptr, re, o * E, 0), buf_ld((T*)nullptr, im, o * E, 0)};
    }
    // element at vo + so, so wave-uniform (scalar offset); the range is not checked against so
    __device__ cx<T> ld(uint32_t vo, uint32_t so, int) const {
        return {buf_ld((T*)nullptr, re, vo * E, so * E), buf_ld((T*)nullptr, im, vo * E, so * E)};
    }
};
template<class T = float>
struct BufW {
    static constexpr uint32_t E = sizeof(T);
    __amdgpu_buffer_rsrc_t re, im;
    __device__ BufW(T* base, int lp) : re(rsrc(base)), im(rsrc(base + ((int64_t)1 << lp))) {}
    // Stores at element offsets >= nfl are dropped by the range check (nfl = poff(lim): exactly
    // the elements i >= lim)
    __device__ BufW(T* base, int lp, uint32_t nfl)
        : re(rsrc(base, nfl * E)), im(rsrc(base + ((int64_t)1 << lp), nfl * E)) {}
    __device__ void st(uint32_t o, cx<T> v, int) const {
        buf_st(v.x, re, o * E, 0);
        buf_st(v.y, im, o * E, 0);
    }
    __device__ void st(uint32_t vo, uint32_t so, cx<T> v, int) const {
        buf_st(v.x, re, vo * E, so * E);
        buf_st(v.y, im, vo * E, so * E);
    }
};
template<bool UNI, class T = float>
using RdView = typename std::conditional<UNI, BufR<T>, CxPtr<T>>::type;
template<bool UNI, class T = float>
using WrView = typename std::conditional<UNI, BufW<T>, CxPtrW<T>>::type;

// LDS copy of the two-level W_N table for four-step pass twiddles, TWN complex entries
// (lo + hi, TWN >= 4); TWN == 2 reads both levels from global memory instead.
template<int TWN, class T = float>
struct PassTw {
    static constexpr bool in_lds = TWN > 2;
    __device__ static void fill(T* t, const vec2<T>* __restrict__ twl, const vec2<T>* __restrict__ twh, int ls,
                                int lg, int tid, int nthreads) {
        if constexpr (in_lds) {
            int nl = 1 << ls, nh = 1 << (lg - ls);
            for (int e = tid; e < nl; e += nthreads) {
                vec2<T> w = twl[e];
                t[e]      = w.x;
                t[nl + e] = w.y;
            }
            for (int h = tid; h < nh; h += nthreads) {
                vec2<T> w          = twh[h];
                t[2 * nl + h]      = w.x;
                t[2 * nl + nh + h] = w.y;
            }
        }
    }
    __device__ static cx<T> get(const T* t, const vec2<T>* __restrict__ twl, const vec2<T>* __restrict__ twh, int ls,
                                int lg, uint32_t e) {
        uint32_t lo = e & ((1u << ls) - 1), hi = e >> ls;
        if constexpr (in_lds) {
            int nl = 1 << ls, nh = 1 << (lg - ls);
            return cmul(cx<T>{t[lo], t[nl + lo]}, cx<T>{t[2 * nl + hi], t[2 * nl + nh + hi]});
        } else {
            vec2<T> a = twl[lo], b = twh[hi];
            return cmul(cx<T>{a.x, a.y}, cx<T>{b.x, b.y});
        }
    }
};
// pass-twiddle table capacity for a transform of 2^lg points (0: none needed)
constexpr int pass_tw_entries(int lg) { return (1 << ((lg + 1) / 2)) + (1 << (lg / 2)); }

template<int LA, class T = float>
constexpr int default_rows() {
    return Plan<LA, false, T>::TA >= 256 ? 1 : 256 / Plan<LA, false, T>::TA;
}

// LDS carving: DATA elements of exchange arrays, a twiddle table of TWF elements and PTW of pass
// twiddles.  The tables go first when behind the data they would end past 64 KiB, the reach of
// the LDS instructions' immediate offsets (each lookup would then need an address addition);
// otherwise the data stays at offset 0.
template<class T, int DATA, int TWF, int PTW>
struct LdsCarve {
    static constexpr bool FIRST = (DATA + TWF + PTW) * (int)sizeof(T) > 65536;
    static constexpr int  DOFF  = FIRST ? TWF + PTW : 0; // exchange arrays
    static constexpr int  TOFF  = FIRST ? 0 : DATA;      // twiddle table, then pass twiddles
};

template<int LA, int G_, class T = float, int PX = 0>
struct RowShape {
    static constexpr int TA = Plan<LA, false, T, PX>::TA;
    static constexpr int G  = G_; // rows per workgroup
    static constexpr int NT = G * TA;
    // odd row stride: rows of a strided access land in distinct banks
    static constexpr int STRIDE = (1 << LA) + ((1 << LA) >> 5) + 1;
    // re and im arrays, or one array shared by both halves of each exchange (Plan::HALF)
    static constexpr int DATA   = (Plan<LA, false, T, PX>::HALF ? 1 : 2) * G * STRIDE;
    __device__ static T* im(T* lre) { return Plan<LA, false, T, PX>::HALF ? lre : lre + DATA / 2; }
};

// Waves per SIMD that the LDS footprint allows (argument of __launch_bounds__).
constexpr int waves_per_simd(int threads, int lds_bytes) {
    int wg = (160 * 1024) / (lds_bytes > 0 ? lds_bytes : 1);
    if (wg > 8)
        wg = 8;
    if (wg < 1)
        wg = 1;
    int w = wg * ((threads + 63) / 64) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// Largest rows that take the row-major tile load / store of row_group (at most 16 threads per
// row), for natural-order and for permuted (bit_reversed / unordered) rows.  Per 2 GiB batch on
// MI355X: natural order FP32 2^1 .. 2^8 0.75-0.92 ms with it against 0.78-2.95 without, FP64
// 2^2 .. 2^5 0.77-0.79 against 1.0-2.0 (FP64 2^7, 2^8: 4-6 % slower with it); bit_reversed FP32
// 2^1 .. 2^5 0.84-1.29 against 1.21-2.43 (2^6 .. 2^8: 1.17-1.33 against 0.80-0.91), FP64
// 2^1 .. 2^7 0.76-0.88 against 0.84-1.42.
constexpr int kTileMaxLog2        = 8;
constexpr int kTileMaxLog2_64     = 5;
constexpr int kTilePermMaxLog2    = 5;
constexpr int kTilePermMaxLog2_64 = 7;

// One workgroup's group of G rows (rows wg*G .. wg*G + G - 1).
// UNI: the rows of a wavefront share one transform (TA >= 64, or all rows of the workgroup lie
// in one transform), so its base address is wave-uniform and buffer addressing applies.
// Element held by local slot s = jt + u*TA of a permuted (bit_reversed / unordered) row.  When a
// thread's slots differ only in their high bits (TA >= 32: q touches the low five bits, which
// belong to jt) this is base(jt) + brev(u), a compile-time offset from one address.
template<int LA, int ORD, class T = float, int PX = 0>
__device__ __forceinline__ int perm_base(int jt) {
    using P           = Plan<LA, false, T, PX>;
    constexpr int LT  = ilog2c(P::TA);
    constexpr int LPP = ilog2c(P::PPT);
    uint32_t      j   = ORD == ORD_UNORD ? uq((uint32_t)jt) : (uint32_t)jt;
    return (int)(brev_bits(j, LT) << LPP);
}
template<int LA, int ORD, class T = float, int PX = 0>
__device__ __forceinline__ int perm_elem(int jt, int u) {
    using P = Plan<LA, false, T, PX>;
    if constexpr (P::TA >= 32)
        return perm_base<LA, ORD, T, PX>(jt) + brevc(u, ilog2c(P::PPT));
    else
        return (int)slot_freq(ORD, (uint32_t)(jt + u * P::TA), LA);
}

template<int LA, int G, int LD, int ST, int TW, bool UNI, int ORD, int PREC = PREC_FP32, class T = float, int PX = 0>
__device__ __forceinline__ void row_group(const RowArgsT<T>& p, int64_t wg, T* lre, T* lim, const T* ltw,
                                          const T* lptw, const MfmaDft16* mf = nullptr) {
    using P           = Plan<LA, false, T, PX>;
    using S           = RowShape<LA, G, T, PX>;
    constexpr int TA  = P::TA;
    constexpr int PPT = P::PPT;
    constexpr int NT  = S::NT;
    constexpr int STR = S::STRIDE;
    constexpr int R0  = P::first_radix;
    constexpr int RL  = P::last_radix;

    const int      t      = threadIdx.x;
    const int      g      = t / TA;
    const int      jt     = t % TA;
    const int64_t  row0   = wg * G;
    // a wavefront holds one row when TA >= 64: keep the row (and its base address) in SGPRs
    const int64_t  row    = TA >= 64 ? row0 + __builtin_amdgcn_readfirstlane(g) : row0 + g;
    const bool     row_ok = row < p.nrows;
    const int64_t  rmask  = ((int64_t)1 << p.rows_lg) - 1;
    const int64_t  bt     = row >> p.rows_lg;
    const uint32_t k2     = (uint32_t)(row & rmask);
    // the row the contiguous loads read (k2 unless the row index is split)
    const uint32_t k2l    = p.rsplit_lg == 0 ? k2
                                             : ((k2 & ((1u << p.rsplit_lg) - 1)) << (p.rows_lg - p.rsplit_lg)) |
                                                   (k2 >> p.rsplit_lg);
    // transform of this thread's row; with UNI the same for the whole wavefront.  Rows past the
    // batch (the last group may be partial) take the group's first row, which always exists:
    // their loads stay inside the source (values masked) and they store nothing.
    const int64_t  btu    = (UNI && TA < 64) || !row_ok ? (row0 >> p.rows_lg) : bt;
    auto           addr   = [&](int b, int o) { return g * STR + b + (b >> 5) + o + (o >> 5); };
    // The permuted loads and stores exchange through LDS with one pad per 2^PP elements,
    // PP = max(LA - 6, 5).  In the permuted order lane jt touches element brev(jt) * PPT + c,
    // and the elements of a wave's 64 lanes differ by multiples of 2^(LA - 6): with this pad
    // they fall in 64 distinct banks from 2^11 on (one pad per 32 gives 2-, 4- and 8-way
    // conflicts at 2^12, 2^13, 2^14; 2^10 keeps it and two-way conflicts, as the row stride is
    // sized for one pad per 32).  Offsets stay carry-free: every offset added to a base is a
    // multiple of 2^PP, or smaller than the room left below the next multiple.
    constexpr int  PP     = LA - 6 > 5 ? LA - 6 : 5;
    auto           paddr  = [&](int b, int o) { return g * STR + b + (b >> PP) + o + (o >> PP); };
    // Row-fastest assignment of the strided passes: thread t takes row ga = t mod G and butterfly
    // index ja = t / G, so the lanes of a wave hold consecutive rows (k2) and a strided load or
    // store of element k1 of rows k2 .. k2 + G - 1 touches G consecutive elements.  The first
    // stage (strided load) or the last stage (strided store) runs in this assignment; the LDS
    // exchange next to it does the transposition.
    constexpr bool ALT_LD = LD == LD_STRIDED && !P::HALF && P::NSTAGES >= 2;
    constexpr bool ALT_ST = ST == ST_STRIDED && !P::HALF && P::NSTAGES >= 2;
    const int      ga     = t % G;
    const int      ja     = t / G;
    auto           addrA  = [&](int b, int o) { return ga * STR + b + (b >> 5) + o + (o >> 5); };
    const StageMap<decltype(addrA)> mapA{ja, addrA};

    // Rows of fewer than 16 threads (2^5 points and below) load and store contiguous rows as a
    // row-major tile: element e of the group's row gg belongs to thread (gg * A + e) mod NT, so a
    // wave covers whole rows with consecutive addresses, and LDS hands the threads their
    // butterflies (a thread of a 2^5 row would otherwise touch 8 bytes of every 128-byte line).
    // It applies when the rows are whole transforms at a constant distance and the group's tile
    // fits 32-bit buffer offsets; the buffer range check drops the rows past the batch.
    constexpr bool F64T  = sizeof(T) == 8;
    constexpr bool TILE  = TA <= 16 && !P::HALF && LA <= (F64T ? kTileMaxLog2_64 : kTileMaxLog2);
    constexpr bool TILEP = TA <= 16 && !P::HALF && LA <= (F64T ? kTilePermMaxLog2_64 : kTilePermMaxLog2);
    constexpr int  TROWS = NT >> LA; // tile rows per pass over the threads
    static_assert(!(TILE || TILEP) || (NT >= P::A && NT % P::A == 0), "a tile pass covers whole rows");
    const int      te    = t & (P::A - 1);
    const int      tg    = t >> LA;
    // A row's footprint in either view (elements poff(0) .. poff(A - 1)); a single row takes it
    // as its distance, so the range check leaves the tile's other rows out.
    auto tile_dist = [&](int lp, int64_t dist) {
        const int64_t foot = (int64_t)poff((uint32_t)(P::A - 1), lp) + 1;
        return p.nrows == 1 ? foot : (dist >= foot ? dist : (int64_t)-1);
    };
    auto tile_ok = [&](const int64_t* offs, int64_t d) {
        return offs == nullptr && p.rows_lg == 0 && d > 0 && (int64_t)G * d * (int64_t)sizeof(T) < 0x7FFFFFFF;
    };
    const uint32_t tile_n = (uint32_t)(p.nrows - row0 < G ? p.nrows - row0 : G); // rows of this group
    // range of the group's tile: its valid rows, up to the last element of the last one
    auto tile_range = [&](int lp, uint32_t d) { return (tile_n - 1) * d + poff((uint32_t)(P::A - 1), lp) + 1; };

    cx<T> v[PPT];
    // ------------------------------------------------------------------ load
    bool loaded = false;
    if constexpr ((LD == LD_CONTIG && TILE) || (LD == LD_PERM && TILEP)) {
        const int64_t sd = tile_dist(p.src_lp, p.src_dist);
        if (tile_ok(p.src_offs, sd)) {
            const uint32_t dist = (uint32_t)sd;
            const BufR<T>  src(p.src + row0 * p.src_dist, p.src_lp, tile_range(p.src_lp, dist));
            const bool     ez = (int64_t)te < p.src_size; // zero extension past src_size
            // a permuted row holds frequency slot_freq(e) at position e
            const int      f  = LD == LD_PERM ? (int)slot_freq(ORD, (uint32_t)te, LA) : te;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                const int gg = tg + u * TROWS;
                cx<T>     x  = src.ld((uint32_t)gg * dist + poff((uint32_t)te, p.src_lp), p.nt_load);
                x            = ez ? x : cx<T>{T(0), T(0)};
                const int o  = gg * STR + f + (f >> 5);
                lre[o]       = x.x;
                lim[o]       = x.y;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int o   = addr(jt + m * TA, r * (P::A / R0));
                    v[m * R0 + r] = cx<T>{lre[o], lim[o]};
                }
            loaded = true;
        }
    }
    if (loaded) {
    } else if constexpr (LD == LD_CONTIG) {
        const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, btu), p.src_lp);
        if (row_ok && (int64_t)(k2l + 1) << LA <= p.src_size) { // whole row inside the source
            if constexpr (P::HALF) {
                // one per-lane offset plus scalar offsets: the 2^14-point plans need the registers
                const uint32_t vo = poff((uint32_t)jt, p.src_lp);
#pragma unroll
                for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        uint32_t c    = (uint32_t)(in_elem<LA, false, T, PX>(jt, m, r) - jt) + (k2l << LA);
                        v[m * R0 + r] = src.ld(vo, poff(c, p.src_lp), p.nt_load);
                    }
            } else {
#pragma unroll
                for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        uint32_t i    = (uint32_t)in_elem<LA, false, T, PX>(jt, m, r) + (k2l << LA);
                        v[m * R0 + r] = src.ld(poff(i, p.src_lp), p.nt_load);
                    }
            }
        } else {
#pragma unroll
            for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    uint32_t i    = (uint32_t)in_elem<LA, false, T, PX>(jt, m, r) + (k2l << LA);
                    bool     ok   = row_ok && i < p.src_size;
                    cx<T>    x    = src.ld(ok ? poff(i, p.src_lp) : 0u, p.nt_load);
                    v[m * R0 + r] = ok ? x : cx<T>{T(0), T(0)};
                }
        }
    } else if constexpr (P::HALF) {
        // one LDS array: gather into registers in slot order, then real parts through LDS into
        // the first stage's layout, then imaginary parts
        auto pos = [&](int u) {
            if constexpr (LD == LD_PERM) {
                if constexpr (TA >= 32)
                    return paddr(perm_base<LA, ORD, T, PX>(jt), brevc(u, ilog2c(PPT)));
                else
                    return paddr(perm_elem<LA, ORD, T, PX>(jt, u), 0);
            } else {
                int idx = t + u * NT;
                int k1  = idx / G;
                return (idx % G) * STR + k1 + (k1 >> 5);
            }
        };
        if constexpr (LD == LD_PERM) {
            const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, btu), p.src_lp);
            uint32_t          base = brev_bits(k2, p.rows_lg) << LA;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                cx<T> x = src.ld(row_ok ? poff(base + jt + u * TA, p.src_lp) : 0u, p.nt_load);
                v[u] = row_ok ? x : cx<T>{T(0), T(0)};
            }
        } else {
            const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, row0 >> p.rows_lg), p.src_lp);
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int     idx = t + u * NT;
                int64_t rw  = row0 + idx % G;
                bool    ok  = rw < p.nrows;
                cx<T>   x   = src.ld(ok ? poff((uint32_t)(rw & rmask) + ((uint32_t)(idx / G) << p.rows_lg), p.src_lp)
                                        : 0u,
                                     p.nt_load);
                v[u]        = ok ? x : cx<T>{T(0), T(0)};
            }
        }
        // the first stage's layout, in the padding that pos() wrote
        auto rd = [&](int b, int o) {
            if constexpr (LD == LD_PERM)
                return paddr(b, o);
            else
                return addr(b, o);
        };
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            lre[pos(u)] = v[u].x;
        __syncthreads();
        stage_read<LA, false, 0, 0, PX>(v, lre, lim, jt, rd);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            lre[pos(u)] = v[u].y;
        __syncthreads();
        stage_read<LA, false, 0, 1, PX>(v, lre, lim, jt, rd);
    } else {
        if constexpr (LD == LD_PERM) {
            // this row's spectrum occupies slots [A*brev(k2), A*brev(k2) + A); local slot s holds
            // element slot_freq(order, s).  All loads first, then the LDS writes: interleaving
            // them let the compiler wait for each load before the next one (2^11: 1.27 ms per
            // 2 GiB against 0.76 for the forward transform).
            const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, btu), p.src_lp);
            uint32_t          base = brev_bits(k2, p.rows_lg) << LA;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                cx<T> x = src.ld(row_ok ? poff(base + jt + u * TA, p.src_lp) : 0u, p.nt_load);
                v[u]    = row_ok ? x : cx<T>{T(0), T(0)};
            }
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int o;
                if constexpr (TA >= 32)
                    o = paddr(perm_base<LA, ORD, T, PX>(jt), brevc(u, ilog2c(PPT)));
                else
                    o = paddr(perm_elem<LA, ORD, T, PX>(jt, u), 0);
                lre[o] = v[u].x;
                lim[o] = v[u].y;
            }
        } else if constexpr (ALT_LD) {
            // LD_STRIDED, row-fastest: element k1 of row k2 sits at slot k2 + (k1 << rows_lg); the
            // stage-0 inputs of this thread's butterflies go straight to registers, G consecutive
            // slots per wave and element (the rows of a workgroup lie in one transform)
            const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, row0 >> p.rows_lg), p.src_lp);
            const int64_t        rw = row0 + ga;
            const bool           ok = rw < p.nrows;
            const uint32_t       kb = (uint32_t)(rw & rmask);
#pragma unroll
            for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const uint32_t k1 = (uint32_t)in_elem<LA, false, T, PX>(ja, m, r);
                    cx<T>          x  = src.ld(ok ? poff(kb + (k1 << p.rows_lg), p.src_lp) : 0u, p.nt_load);
                    v[m * R0 + r]     = ok ? x : cx<T>{T(0), T(0)};
                }
        } else { // LD_STRIDED: element k1 of row k2 sits at slot k2 + (k1 << rows_lg)
            // (the rows of a workgroup lie in one transform: G divides the rows per transform)
            const RdView<UNI, T> src(tbase(p.src, p.src_offs, p.src_dist, row0 >> p.rows_lg), p.src_lp);
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int     idx = t + u * NT;
                int     gg  = idx % G;
                int     k1  = idx / G;
                int64_t rw  = row0 + gg;
                bool    ok  = rw < p.nrows;
                cx<T>   x   = src.ld(ok ? poff((uint32_t)(rw & rmask) + ((uint32_t)k1 << p.rows_lg), p.src_lp) : 0u,
                                     p.nt_load);
                x           = ok ? x : cx<T>{T(0), T(0)};
                int o  = gg * STR + k1 + (k1 >> 5);
                lre[o] = x.x;
                lim[o] = x.y;
            }
        }
        if constexpr (!ALT_LD) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    int o = LD == LD_PERM ? paddr(jt + m * TA, r * (P::A / R0)) : addr(jt + m * TA, r * (P::A / R0));
                    v[m * R0 + r] = cx<T>{lre[o], lim[o]};
                }
        }
    }
    if (p.conj_in) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    // ------------------------------------------------------------------ transform
    // (the barrier before the first stage's LDS writes also orders them after the reads above)
    // (jt = t % TA: butterfly index jt of a transform that fits a wave sits in lane jt % 64)
    using AA = decltype(addrA);
    if constexpr (ALT_LD)
        stockham<LA, false, decltype(addr), PREC, T, PX, true, false, AA, AA, true>(v, lre, lim, jt, addr, ltw, mf,
                                                                                  &mapA, &mapA);
    else if constexpr (ALT_ST)
        stockham<LA, false, decltype(addr), PREC, T, PX, false, true, AA, AA, true>(v, lre, lim, jt, addr, ltw, mf,
                                                                                  &mapA, &mapA);
    else
        stockham<LA, false, decltype(addr), PREC, T, PX, false, false, AA, AA, true>(v, lre, lim, jt, addr, ltw, mf);
    // the thread's outputs: butterfly index and row of the last stage's assignment
    const int      jo  = ALT_ST ? ja : jt;
    const uint32_t ko2 = ALT_ST ? (uint32_t)((row0 + ga) & rmask) : k2;

    if constexpr (TW > 0) {
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t e    = (uint32_t)out_elem<LA, false, T, PX>(jo, m, r) * ko2;
                v[m * RL + r] = cmul(v[m * RL + r], PassTw<TW, T>::get(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, e));
                __builtin_amdgcn_sched_barrier(0); // one element's table reads in flight at a time
            }
    }
    if (p.scale != T(1)) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            v[q].x *= p.scale;
            v[q].y *= p.scale;
        }
    }
    if (p.conj_out) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    // ------------------------------------------------------------------ store
    bool stored = false;
    if constexpr ((ST == ST_CONTIG && TILE) || (ST == ST_PERM && TILEP)) {
        const int64_t dd = tile_dist(p.dst_lp, p.dst_dist);
        if (tile_ok(p.dst_offs, dd)) {
            __syncthreads(); // the last stage's LDS reads are done
#pragma unroll
            for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int o = addr(jt + m * TA, r * (P::A / RL));
                    lre[o]      = v[m * RL + r].x;
                    lim[o]      = v[m * RL + r].y;
                }
            __syncthreads();
            const uint32_t dist = (uint32_t)dd;
            const BufW<T>  dst(p.dst + row0 * p.dst_dist, p.dst_lp, tile_range(p.dst_lp, dist));
            const int      f = ST == ST_PERM ? (int)slot_freq(ORD, (uint32_t)opaque_after_memory(te), LA) : te;
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                const int gg = tg + u * TROWS;
                const int o  = gg * STR + f + (f >> 5);
                dst.st((uint32_t)gg * dist + poff((uint32_t)te, p.dst_lp), cx<T>{lre[o], lim[o]}, p.nt_store);
            }
            stored = true;
        }
    }
    if (stored) {
    } else if constexpr (ST == ST_CONTIG) {
        if (row_ok) {
            const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, btu), p.dst_lp);
            if constexpr (P::HALF) {
                const uint32_t vo = poff((uint32_t)jt, p.dst_lp);
#pragma unroll
                for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                    for (int r = 0; r < RL; ++r) {
                        uint32_t c = (uint32_t)(out_elem<LA, false, T, PX>(jt, m, r) - jt) + (k2 << LA);
                        dst.st(vo, poff(c, p.dst_lp), v[m * RL + r], p.nt_store);
                    }
            } else {
#pragma unroll
                for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                    for (int r = 0; r < RL; ++r) {
                        uint32_t i = (uint32_t)out_elem<LA, false, T, PX>(jt, m, r) + (k2 << LA);
                        dst.st(poff(i, p.dst_lp), v[m * RL + r], p.nt_store);
                    }
            }
        }
    } else if constexpr (P::HALF) {
        // one LDS array: real parts into slot order, then imaginary parts, then the stores
        auto pos = [&](int u) {
            if constexpr (ST == ST_PERM) {
                if constexpr (TA >= 32)
                    return paddr(perm_base<LA, ORD, T, PX>(jt), brevc(u, ilog2c(PPT)));
                else
                    return paddr(perm_elem<LA, ORD, T, PX>(jt, u), 0);
            } else {
                int idx = t + u * NT;
                int k1  = idx / G;
                return (idx % G) * STR + k1 + (k1 >> 5);
            }
        };
        auto put = [&](auto comp) {
#pragma unroll
            for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int o = ST == ST_PERM ? paddr(jt + m * TA, r * (P::A / RL)) : addr(jt + m * TA, r * (P::A / RL));
                    lre[o]      = decltype(comp)::value ? v[m * RL + r].y : v[m * RL + r].x;
                }
        };
        __syncthreads(); // the last stage's LDS reads are done
        put(std::false_type{});
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            v[u].x = lre[pos(u)];
        __syncthreads();
        put(std::true_type{}); // imaginary parts are still in the stage layout
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            v[u].y = lre[pos(u)];
        if constexpr (ST == ST_PERM) {
            if (row_ok) {
                const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, btu), p.dst_lp);
                uint32_t          base = brev_bits(k2, p.rows_lg) << LA;
#pragma unroll
                for (int u = 0; u < PPT; ++u)
                    dst.st(poff(base + jt + u * TA, p.dst_lp), v[u], p.nt_store);
            }
        } else {
            const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, row0 >> p.rows_lg), p.dst_lp);
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int     idx = t + u * NT;
                int64_t rw  = row0 + idx % G;
                if (rw < p.nrows)
                    dst.st(poff((uint32_t)(rw & rmask) + ((uint32_t)(idx / G) << p.rows_lg), p.dst_lp), v[u],
                           p.nt_store);
            }
        }
    } else if constexpr (ALT_ST) {
        // ST_STRIDED, row-fastest: element k1 of row k2 goes to slot k2 + (k1 << rows_lg), G
        // consecutive slots per wave and element, straight from registers
        const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, row0 >> p.rows_lg), p.dst_lp);
        const int64_t        rw = row0 + ga;
        if (rw < p.nrows) {
#pragma unroll
            for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const uint32_t k1 = (uint32_t)out_elem<LA, false, T, PX>(ja, m, r);
                    dst.st(poff(ko2 + (k1 << p.rows_lg), p.dst_lp), v[m * RL + r], p.nt_store);
                }
        }
    } else {
        __syncthreads(); // the last stage's LDS reads are done
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                int o  = ST == ST_PERM ? paddr(jt + m * TA, r * (P::A / RL)) : addr(jt + m * TA, r * (P::A / RL));
                lre[o] = v[m * RL + r].x;
                lim[o] = v[m * RL + r].y;
            }
        __syncthreads();
        if constexpr (ST == ST_PERM) {
            if (row_ok) {
                const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, btu), p.dst_lp);
                uint32_t          base = brev_bits(k2, p.rows_lg) << LA;
#pragma unroll
                for (int u = 0; u < PPT; ++u) {
                    int s = jt + u * TA;
                    int o;
                    if constexpr (TA >= 32)
                        o = paddr(perm_base<LA, ORD, T, PX>(jt), brevc(u, ilog2c(PPT)));
                    else
                        o = paddr(perm_elem<LA, ORD, T, PX>(jt, u), 0);
                    dst.st(poff(base + s, p.dst_lp), cx<T>{lre[o], lim[o]}, p.nt_store);
                }
            }
        } else { // ST_STRIDED: element k1 of row k2 goes to slot k2 + (k1 << rows_lg)
            const WrView<UNI, T> dst(tbase(p.dst, p.dst_offs, p.dst_dist, row0 >> p.rows_lg), p.dst_lp);
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                int     idx = t + u * NT;
                int     gg  = idx % G;
                int     k1  = idx / G;
                int64_t rw  = row0 + gg;
                if (rw < p.nrows) {
                    int o = gg * STR + k1 + (k1 >> 5);
                    dst.st(poff((uint32_t)(rw & rmask) + ((uint32_t)k1 << p.rows_lg), p.dst_lp), cx<T>{lre[o], lim[o]},
                           p.nt_store);
                }
            }
        }
    }
}

// LDS elements (of T) of a row workgroup
template<int LA, int G, int TW, class T = float>
constexpr int row_lds_floats() {
    return RowShape<LA, G, T>::DATA + TwTable<LA, T>::FLOATS + 2 * TW;
}

// Grid-stride over row groups: gridDim.x may be smaller than the number of groups.
// The matrix-core variants keep the split operands in registers: they run at half the waves,
// but at least two, so the register budget (256) never pushes accumulators into AGPRs.
constexpr int mfma_waves(int w, int prec) { return prec == PREC_FP32 ? w : (w > 4 ? w / 2 : 2); }

template<int LA, int G, int LD, int ST, int TW, bool UNI, int ORD, int PREC = PREC_FP32, class T = float>
__global__ __launch_bounds__((G * Plan<LA, false, T>::TA),
                             mfma_waves(waves_per_simd(G * Plan<LA, false, T>::TA,
                                                       (int)sizeof(T) * row_lds_floats<LA, G, TW, T>()),
                                        PREC))
void k_row(RowArgsT<T> p) {
    constexpr int NT = G * Plan<LA, false, T>::TA;
    __shared__ T  lds[row_lds_floats<LA, G, TW, T>()];
    using LC           = LdsCarve<T, RowShape<LA, G, T>::DATA, TwTable<LA, T>::FLOATS, 2 * TW>;
    T*            ltw  = lds + LC::TOFF;
    T*            lptw = ltw + TwTable<LA, T>::FLOATS;
    T*            lre  = lds + LC::DOFF;
    T*            lim  = RowShape<LA, G, T>::im(lre);
    TwTable<LA, T>::fill(ltw, p.twA, threadIdx.x, NT);
    if constexpr (TW > 0)
        PassTw<TW, T>::fill(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, threadIdx.x, NT);
    __syncthreads();
    MfmaDft16 mf;
    if constexpr (PREC != PREC_FP32)
        mf.template init<PREC>();
    const int64_t nwg = (p.nrows + G - 1) / G;
    for (int64_t base = 0; base < nwg; base += gridDim.x) {
        const int64_t wg = xcd_item(base);
        if (wg >= nwg)
            break; // uniform over the workgroup; later rounds have no items for it either
        row_group<LA, G, LD, ST, TW, UNI, ORD, PREC, T>(p, wg, lre, lim, ltw, lptw, &mf);
        __syncthreads(); // LDS is reused by the next group
    }
}

// ---------------------------------------------------------------------------------------------
// k_dif: one-pass transform of N = R * M points (R = 2^LR, M = 2^LM <= 16384) with no HBM
// intermediate.  With n = n2 + M*n1 and k = R*k2 + k1 (one decimation-in-frequency step of
// radix R):
//   X[R*k2 + k1] = FFT_M(y_k1)[k2],   y_k1[n2] = W_N^(n2*k1) * sum_{n1<R} x[n2 + M*n1] W_R^(n1*k1).
// Workgroup (b, k1) reads all of transform b (R segments of M points), forms y_k1 in registers
// and runs the M-point Stockham transform in LDS.  The R workgroups of one transform read the
// same bytes: they are numbered w = 8*(R*t + k1) + x for transform b = 8*t + x, so the
// round-robin placement of workgroups on the 8 XCDs (w mod 8) puts them on one XCD, dispatched
// back to back, and the repeated reads are served by that XCD's L2.  (Placement only affects
// speed, never results.)  Stores: natural order puts k2 of workgroup k1 at R*k2 + k1 (the R
// workgroups fill interleaved floats of each cache line; a line evicted before all R have written
// costs a second HBM write); bit_reversed / unordered give workgroup k1 the contiguous block
// brev_R(k1)*M + [0, M), permuted inside.  Out of place only: every workgroup reads the whole
// transform.
template<int LM, int LR>
constexpr int dif_lds_floats() {
    return RowShape<LM, 1>::DATA + TwTable<LM>::FLOATS + 2 * pass_tw_entries(LM + LR);
}
// Plan options of k_dif: 32 points per thread for M = 2^14 (512 threads; 1.32 against 1.43 ms
// per 2 GiB at N = 2^15 with 64), else the default rule; twiddle exponents by addition chains
// (one sub-transform per workgroup: no loop to hoist the twiddle addresses out of).
template<int LM>
constexpr int dif_ppt() {
    return kTwChain | (LM >= 14 ? 32 : 0);
}

template<int LM, int LR, int ORD, int PREC = PREC_FP32>
__global__ __launch_bounds__((Plan<LM, false, float, dif_ppt<LM>()>::TA),
                             mfma_waves(waves_per_simd(Plan<LM, false, float, dif_ppt<LM>()>::TA,
                                                       4 * dif_lds_floats<LM, LR>()),
                                        PREC))
void k_dif(RowArgs p) {
    constexpr int PX  = dif_ppt<LM>();
    using P           = Plan<LM, false, float, PX>;
    constexpr int TA  = P::TA;
    constexpr int PPT = P::PPT;
    constexpr int R0  = P::first_radix;
    constexpr int RL  = P::last_radix;
    constexpr int M   = 1 << LM;
    constexpr int R   = 1 << LR;
    constexpr int TWN = pass_tw_entries(LM + LR);
    static_assert(LM >= 5 && TA >= 32, "k_dif needs the block-permuted store layout (TA >= 32)");
    static_assert(!P::HALF || ORD == ORD_NORMAL, "permuted k_dif stores stage through separate re / im arrays");
    __shared__ float lds[dif_lds_floats<LM, LR>()];
    using LC              = LdsCarve<float, RowShape<LM, 1>::DATA, TwTable<LM>::FLOATS, 2 * TWN>;
    float*           ltw  = lds + LC::TOFF;
    float*           lptw = ltw + TwTable<LM>::FLOATS;
    float*           lre  = lds + LC::DOFF;
    float*           lim  = RowShape<LM, 1>::im(lre);

    const uint32_t w   = blockIdx.x;
    const int      k1  = (int)((w >> 3) & (R - 1));
    const int64_t  b   = (int64_t)(w >> (3 + LR)) * 8 + (w & 7);
    if (b >= p.nrows)
        return; // uniform over the workgroup
    const int jt = threadIdx.x;
    TwTable<LM>::fill(ltw, p.twA, jt, TA);
    PassTw<TWN>::fill(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, jt, TA);
    MfmaDft16 mf;
    if constexpr (PREC != PREC_FP32)
        mf.template init<PREC>();
    auto addr = [&](int bb, int o) { return bb + (bb >> 5) + o + (o >> 5); };

    // ------------------------------------------------------------------ load + radix-R step
    cf v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q)
        v[q] = cf{0.f, 0.f};
    {
        const float*   tb  = tbase(p.src, p.src_offs, p.src_dist, b);
        const float    cs  = p.conj_in ? -1.f : 1.f;
        const uint32_t lim = p.src_size < (int64_t)R * M ? (uint32_t)p.src_size : (uint32_t)(R * M);
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) {
            // W_R^(n1*k1) = W_M^(n1*k1*M/R), a uniform read of the W_M table
            const float2   w2 = p.twA[((n1 * k1) & (R - 1)) << (LM - LR)];
            const cf       wr = cf{w2.x, w2.y};
            const uint32_t s0 = (uint32_t)n1 << LM;
            // segment n1 starts at poff(n1*M) (poff(n2 + n1*M) = poff(n2) + poff(n1*M) for every
            // power-of-two pack); elements past src_size read as zero through the range check
            const uint32_t nseg = lim > s0 ? (lim - s0 < (uint32_t)M ? lim - s0 : (uint32_t)M) : 0u;
            const BufR<>   seg(tb + poff(s0, p.src_lp), p.src_lp, poff(nseg, p.src_lp));
            // the per-lane offset of segment n1 > 0 consumes the sums so far (an opaque zero is
            // added), so one segment's loads are in registers at a time beside v
            int jv = jt;
            if (n1 > 0)
                asm volatile("" : "+v"(jv) : "v"(v[0].x), "v"(v[PPT - 1].y));
            const uint32_t vo = poff((uint32_t)jv, p.src_lp);
#pragma unroll
            for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    uint32_t c = (uint32_t)(in_elem<LM, false, float, PX>(jt, m, r) - jt);
                    cf       x = seg.ld(vo + poff(c, p.src_lp), 0);
                    x.y *= cs;
                    if (n1 == 0)
                        v[m * R0 + r] = x;
                    else
                        v[m * R0 + r] = cadd(v[m * R0 + r], cmul(x, wr));
                }
        }
    }
    __syncthreads(); // twiddle tables are in LDS
    if (k1 != 0) {
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                uint32_t e    = (uint32_t)in_elem<LM, false, float, PX>(jt, m, r) * (uint32_t)k1;
                v[m * R0 + r] = cmul(v[m * R0 + r], PassTw<TWN>::get(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, e));
            }
    }
    // ------------------------------------------------------------------ M-point transform
    stockham<LM, false, decltype(addr), PREC, float, PX>(v, lre, lim, jt, addr, ltw, &mf);
    const float sc = p.scale;
    const float so = p.conj_out ? -sc : sc;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        v[q].x *= sc;
        v[q].y *= so;
    }
    // ------------------------------------------------------------------ store
    const BufW<> dst(tbase(p.dst, p.dst_offs, p.dst_dist, b), p.dst_lp);
    if constexpr (ORD == ORD_NORMAL) {
        // element (jt + c) * R + k1 with c a multiple of TA: poff of it is
        // poff(jt * R + k1) + poff(c * R) (the same splitting at the multiple R * TA)
        const int      jt = opaque_after_memory(threadIdx.x);
        const uint32_t vo = poff(((uint32_t)jt << LR) + (uint32_t)k1, p.dst_lp);
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m) {
            const int lp = opaque_s(p.dst_lp);
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t c = (uint32_t)(out_elem<LM, false, float, PX>(jt, m, r) - jt);
                dst.st(vo, poff(c << LR, lp), v[m * RL + r], p.nt_store);
            }
        }
    } else {
        // the permuted exchange's padding (one pad per 2^(LM - 6) elements: conflict-free in
        // both orders, as in row_group)
        constexpr int PP    = LM - 6 > 5 ? LM - 6 : 5;
        auto          paddr = [&](int bb, int o) { return bb + (bb >> PP) + o + (o >> PP); };
        __syncthreads(); // the last stage's LDS reads are done
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                int o  = paddr(jt + m * TA, r * (M / RL));
                lre[o] = v[m * RL + r].x;
                lim[o] = v[m * RL + r].y;
            }
        __syncthreads();
        const uint32_t base = brev_bits((uint32_t)k1, LR) << LM;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            int s = jt + u * TA;
            int o = paddr(perm_base<LM, ORD>(jt), brevc(u, ilog2c(PPT)));
            dst.st(poff(base + s, p.dst_lp), cf{lre[o], lim[o]}, p.nt_store);
        }
    }
}

template<int LA, int C, int TW, int PREC = PREC_FP32, class T = float, int PX = 0>
__device__ __forceinline__ void col_tile(const ColArgsT<T>& p, int64_t tile, T* lre, T* lim, const T* ltw,
                                         const T* lptw, const MfmaDft16* mf = nullptr) {
    using P           = Plan<LA, false, T, PX>;
    constexpr int PPT = P::PPT;
    constexpr int R0  = P::first_radix;
    constexpr int RL  = P::last_radix;

    const int      t    = threadIdx.x;
    const int      c    = t % C;
    const int      jt   = t / C;
    const int      ctl  = p.n1_lg - ilog2c(C); // log2(column tiles per transform)
    const int64_t  bt   = tile >> ctl;
    const uint32_t n1   = (uint32_t)((tile & (((int64_t)1 << ctl) - 1)) * C) + c;
    auto           addr = [&](int b, int o) { return (b + o) * C + c; };
    const BufR<T>  src(tbase(p.src, p.src_offs, p.src_dist, bt), p.src_lp);
    const BufW<T>  dst(tbase(p.dst, p.dst_offs, p.dst_dist, bt), p.dst_lp);

    // Element e = jt + c of the column sits at n1 + (e << n1_lg).  When poff adds over multiples
    // of 2^n1_lg (pack at most 2^n1_lg, or at least the whole transform), each access is one
    // per-lane offset poff(n1 + (jt << n1_lg)) plus the scalar poff(c << n1_lg).
    auto additive = [&](int lp) { return lp <= p.n1_lg || lp >= p.n1_lg + LA; };
    cx<T> v[PPT];
    if (((int64_t)1 << (p.n1_lg + LA)) <= p.src_size && additive(p.src_lp)) {
        const uint32_t vo = poff(n1 + ((uint32_t)jt << p.n1_lg), p.src_lp);
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const uint32_t c = (uint32_t)(in_elem<LA, false, T, PX>(jt, m, r) - jt);
                v[m * R0 + r]    = src.ld(vo, poff(c << p.n1_lg, p.src_lp), p.nt_load);
            }
    } else if (((int64_t)1 << (p.n1_lg + LA)) <= p.src_size) { // the whole transform is inside
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                uint32_t i    = n1 + ((uint32_t)in_elem<LA, false, T, PX>(jt, m, r) << p.n1_lg);
                v[m * R0 + r] = src.ld(poff(i, p.src_lp), p.nt_load);
            }
    } else {
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                uint32_t i    = n1 + ((uint32_t)in_elem<LA, false, T, PX>(jt, m, r) << p.n1_lg);
                bool     ok   = i < p.src_size;
                cx<T>    x    = src.ld(ok ? poff(i, p.src_lp) : 0u, p.nt_load);
                v[m * R0 + r] = ok ? x : cx<T>{T(0), T(0)};
            }
    }
    if (p.conj_in) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    stockham<LA, false, decltype(addr), PREC, T, PX>(v, lre, lim, jt, addr, ltw, mf);
    if constexpr (TW > 0) {
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t e    = n1 * (uint32_t)out_elem<LA, false, T, PX>(jt, m, r);
                v[m * RL + r] = cmul(v[m * RL + r], PassTw<TW, T>::get(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, e));
                __builtin_amdgcn_sched_barrier(0); // one element's table reads in flight at a time
            }
    }
    if (p.scale != T(1)) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            v[q].x *= p.scale;
            v[q].y *= p.scale;
        }
    }
    if (p.conj_out) {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            v[q].y = -v[q].y;
    }
    if (additive(p.dst_lp)) {
        const uint32_t vo = poff(n1 + ((uint32_t)opaque_after_memory(jt) << p.n1_lg), p.dst_lp);
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const uint32_t c = (uint32_t)(out_elem<LA, false, T, PX>(jt, m, r) - jt);
                dst.st(vo, poff(c << p.n1_lg, p.dst_lp), v[m * RL + r], p.nt_store);
            }
    } else {
#pragma unroll
        for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                uint32_t i = n1 + ((uint32_t)out_elem<LA, false, T, PX>(jt, m, r) << p.n1_lg);
                dst.st(poff(i, p.dst_lp), v[m * RL + r], p.nt_store);
            }
    }
}

template<int LA, int C, int TW, class T = float, int PX = 0>
constexpr int col_lds_floats() {
    return (Plan<LA, false, T, PX>::HALF ? 1 : 2) * (1 << LA) * C + TwTable<LA, T>::FLOATS + 2 * TW;
}

// PX: plan options of the columns (points per thread, one-array exchange).
template<int LA, int C, int TW, int PREC = PREC_FP32, class T = float, int PX = 0>
__global__ __launch_bounds__((C * Plan<LA, false, T, PX>::TA),
                             mfma_waves(waves_per_simd(C * Plan<LA, false, T, PX>::TA,
                                                       (int)sizeof(T) * col_lds_floats<LA, C, TW, T, PX>()),
                                        PREC))
void k_col(ColArgsT<T> p) {
    using P            = Plan<LA, false, T, PX>;
    constexpr int NT   = C * P::TA;
    constexpr int DATA = (P::HALF ? 1 : 2) * (1 << LA) * C;
    __shared__ T  lds[col_lds_floats<LA, C, TW, T, PX>()];
    using LC           = LdsCarve<T, DATA, TwTable<LA, T>::FLOATS, 2 * TW>;
    T*            ltw  = lds + LC::TOFF;
    T*            lptw = ltw + TwTable<LA, T>::FLOATS;
    T*            lre  = lds + LC::DOFF;
    T*            lim  = P::HALF ? lre : lre + DATA / 2;
    TwTable<LA, T>::fill(ltw, p.twA, threadIdx.x, NT);
    if constexpr (TW > 0)
        PassTw<TW, T>::fill(lptw, p.twl, p.twh, p.tw_ls, p.tw_lg, threadIdx.x, NT);
    __syncthreads();
    MfmaDft16 mf;
    if constexpr (PREC != PREC_FP32)
        mf.template init<PREC>();
    for (int64_t base = 0; base < p.ntiles; base += gridDim.x) {
        const int64_t tile = xcd_item(base);
        if (tile >= p.ntiles)
            break; // uniform over the workgroup; later rounds have no items for it either
        col_tile<LA, C, TW, PREC, T, PX>(p, tile, lre, lim, ltw, lptw, &mf);
        __syncthreads(); // LDS is reused by the next tile
    }
}

// t[e] = W_M^(e*stride) = exp(-2*pi*i*e*stride/M) for e < count, computed in FP64, rounded to FP32.
// ---------------------------------------------------------------------------------------------
// k_across: FFTs across a range of vectors (the transform the reference's fft_unit_par,
// fft.hpp:2168-2572, computes): for every element index k < len,
//   dst_j[k] = sum_i src_i[k] W_A^(i j),  i, j < A = 2^LA vectors,
// natural order in j.  Vector i is an arbitrary packed vector (pointer table, pack 2^src_lp).  A
// workgroup takes C consecutive elements of all A vectors: a wavefront reads C consecutive floats
// of one vector at a time, and the A-point transforms run down the columns of an LDS tile.
template<class T = float>
struct AcrossArgsT {
    const T* const* src; // A device pointers
    T* const*       dst;
    int64_t         len;
    int             src_lp, dst_lp;
    int             conj_in, conj_out;
    T               scale;
    const vec2<T>*  twA;
};
using AcrossArgs = AcrossArgsT<float>;
template<int LA, int C, class T = float>
constexpr int across_lds_floats() {
    return 2 * (1 << LA) * C + TwTable<LA, T>::FLOATS;
}
template<int LA, int C, class T = float>
__global__ __launch_bounds__((C * Plan<LA, false, T>::TA)) void k_across(AcrossArgsT<T> p) {
    using P            = Plan<LA, false, T>;
    constexpr int PPT  = P::PPT;
    constexpr int R0   = P::first_radix;
    constexpr int RL   = P::last_radix;
    constexpr int NT   = C * P::TA;
    constexpr int DATA = 2 * (1 << LA) * C;
    static_assert(!P::HALF, "across tiles keep separate re / im arrays");
    __shared__ T lds[across_lds_floats<LA, C, T>()];
    using LC         = LdsCarve<T, DATA, TwTable<LA, T>::FLOATS, 0>;
    T*           ltw = lds + LC::TOFF;
    T*           lre = lds + LC::DOFF;
    T*           lim = lre + DATA / 2;
    TwTable<LA, T>::fill(ltw, p.twA, threadIdx.x, NT);
    __syncthreads();
    const int64_t ntiles = (p.len + C - 1) / C;
    const int     c      = threadIdx.x % C;
    const int     jt     = threadIdx.x / C;
    auto          addr   = [&](int b, int o) { return (b + o) * C + c; };
    const T       ci     = p.conj_in ? T(-1) : T(1);
    const T       co     = p.conj_out ? -p.scale : p.scale;
    for (int64_t base = 0; base < ntiles; base += gridDim.x) {
        const int64_t tile = xcd_item(base);
        if (tile >= ntiles)
            break; // uniform over the workgroup
        const int64_t k  = tile * C + c;
        const bool    ok = k < p.len;
        cx<T>         v[PPT];
#pragma unroll
        for (int m = 0; m < PPT / R0; ++m)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const T*      vec = p.src[in_elem<LA, false, T>(jt, m, r)];
                const int64_t o   = pidx(ok ? k : 0, p.src_lp);
                cx<T>         x   = {vec[o], vec[o + ((int64_t)1 << p.src_lp)]};
                v[m * R0 + r]     = ok ? cx<T>{x.x, ci * x.y} : cx<T>{T(0), T(0)};
            }
        stockham<LA, false, decltype(addr), PREC_FP32, T>(v, lre, lim, jt, addr, ltw);
        if (ok) {
            const int64_t o = pidx(k, p.dst_lp);
#pragma unroll
            for (int m = 0; m < PPT / RL; ++m)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    T*          vec                    = p.dst[out_elem<LA, false, T>(jt, m, r)];
                    const cx<T> y                      = v[m * RL + r];
                    vec[o]                             = y.x * p.scale;
                    vec[o + ((int64_t)1 << p.dst_lp)]  = y.y * co;
                }
        }
        __syncthreads(); // LDS is reused by the next tile
    }
}

template<class T>
__global__ void k_twiddles(vec2<T>* t, int64_t M, int64_t stride, int64_t count) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count)
        return;
    int64_t k = (e * stride) % M;
    double  s, c;
    sincospi(-2.0 * (double)k / (double)M, &s, &c);
    t[e] = vec2<T>{(T)c, (T)s};
}

// Seeded inputs: output k of splitmix64(seed) is mix(seed + (k+1)*gamma), so element i uses
// outputs 2i and 2i+1: uniform [-1,1) re/im, or Box-Muller normal(0, sigma^2).
__device__ __forceinline__ uint64_t smix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
template<class T>
__global__ void k_generate(T* dst, int lp, int64_t n, uint64_t seed, int dist, double sigma) {
    const uint64_t gam = 0x9E3779B97F4A7C15ULL;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double u1 = (double)(smix(seed + (uint64_t)(2 * i + 1) * gam) >> 11) * (1.0 / 9007199254740992.0);
        double u2 = (double)(smix(seed + (uint64_t)(2 * i + 2) * gam) >> 11) * (1.0 / 9007199254740992.0);
        T      re, im;
        if (dist == 0) {
            re = (T)(2.0 * u1 - 1.0);
            im = (T)(2.0 * u2 - 1.0);
        } else {
            double r = sqrt(-2.0 * log(1.0 - u1));
            double s, c;
            sincospi(2.0 * u2, &s, &c);
            re = (T)(sigma * r * c);
            im = (T)(sigma * r * s);
        }
        int64_t o                   = i + ((i >> lp) << lp);
        dst[o]                      = re;
        dst[o + ((int64_t)1 << lp)] = im;
    }
}

} // namespace pcxk
