// k_misc.hip -- the persistent uneven-batch queue kernel (one launch for every transform of
// 2^1 .. 2^20 points) and the vector-expression kernel.
#include "dispatch.hpp"

namespace pcxk {

// LDS of the largest item (floats)
template<int LA>
constexpr int queue_single_lds() {
    constexpr int PX = q_ppt(LA);
    constexpr int G  = kQueueThreads / Plan<LA, false, float, PX>::TA;
    return RowShape<LA, G, float, PX>::DATA + TwTable<LA>::FLOATS;
}
template<int LG>
constexpr int queue_four_lds() {
    constexpr int N1L = q_n1(LG), N2L = LG - N1L, C = 16384 >> N2L, G = 16384 >> N1L;
    constexpr int cl  = 2 * (1 << N2L) * C + TwTable<N2L>::FLOATS + 2 * pass_tw_entries(LG);
    constexpr int rl  = RowShape<N1L, G>::DATA + TwTable<N1L>::FLOATS + 2 * pass_tw_entries(LG);
    return cl > rl ? cl : rl;
}
template<int... L>
constexpr int queue_single_max(std::integer_sequence<int, L...>) {
    int m = 0;
    ((m = m > queue_single_lds<L + 1>() ? m : queue_single_lds<L + 1>()), ...);
    return m;
}
template<int... L>
constexpr int queue_four_max(std::integer_sequence<int, L...>) {
    int m = 0;
    ((m = m > queue_four_lds<L + kQueueMinFour>() ? m : queue_four_lds<L + kQueueMinFour>()), ...);
    return m;
}
constexpr int kQueueLdsSingle = queue_single_max(std::make_integer_sequence<int, kQueueMaxSingle>{});
constexpr int kQueueLdsFour   = queue_four_max(std::make_integer_sequence<int, kQueueMaxLog2 - kQueueMinFour + 1>{});
constexpr int kQueueLds       = kQueueLdsSingle > kQueueLdsFour ? kQueueLdsSingle : kQueueLdsFour;
static_assert(4 * kQueueLds <= 160 * 1024, "the largest item fits one CU's LDS");

// The queue kernel's LDS, at file scope: the item functions are compiled separately (each with
// its own register budget) and address it directly as LDS.
__shared__ float g_qlds[kQueueLds];

// Values an item function receives arrive in VGPRs (function arguments); these make the
// workgroup-uniform ones scalar again, so addresses and loop bounds stay in SGPRs.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint64_t u  = (uint64_t)v;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template<class T>
__device__ __forceinline__ T* uni(T* p) {
    return reinterpret_cast<T*>(uni((int64_t)reinterpret_cast<uintptr_t>(p)));
}
__device__ __forceinline__ QItem uni(const QItem& it) { return {uni(it.kind), uni(it.la), uni(it.inv), uni(it.first), uni(it.count)}; }

// ---- single-pass item: `count` transforms of 2^LA points, one row of the workgroup each
template<int LA>
__device__ __noinline__ void queue_single(const QueueArgs& q, const QItem it_) {
    float*      lds = g_qlds;
    const QItem it  = uni(it_);
    constexpr int PX = q_ppt(LA);
    constexpr int TA = Plan<LA, false, float, PX>::TA;
    constexpr int G  = kQueueThreads / TA;
    static_assert(G * TA == kQueueThreads, "single-pass items fill the workgroup");
    using S = RowShape<LA, G, float, PX>;
    RowArgs p{};
    p.src      = uni(q.src);
    p.dst      = uni(q.dst);
    p.src_offs = uni(q.src_offs) + it.first;
    p.dst_offs = uni(q.dst_offs) + it.first;
    p.src_size = (int64_t)1 << LA;
    p.nrows    = it.count;
    p.src_lp = p.dst_lp = LA;
    p.rows_lg           = 0;
    p.order             = 0;
    p.conj_in = p.conj_out = it.inv;
    p.scale                = (it.inv && uni(q.normalized)) ? 1.0f / (float)(1 << LA) : 1.0f;
    p.twA                  = uni(q.tw) + ((1 << LA) - 2);
    using LC               = LdsCarve<float, S::DATA, TwTable<LA>::FLOATS, 0>;
    float* ltw             = lds + LC::TOFF;
    float* lre             = lds + LC::DOFF;
    TwTable<LA>::fill(ltw, p.twA, threadIdx.x, kQueueThreads);
    __syncthreads();
    row_group<LA, G, LD_CONTIG, ST_CONTIG, 0, (TA >= 64), 0, PREC_FP32, float, PX>(p, 0, lre, S::im(lre), ltw, nullptr);
}

// ---- four-step item: one transform of 2^LG points through this workgroup's intermediate.  Each
// pass is its own function (its own register budget); the workgroup runs the first pass over all
// tiles, then the second: the intermediate it wrote is read back by the same workgroup.
template<int LG>
struct QFour {
    static constexpr int N1L = q_n1(LG), N2L = LG - N1L;
    static constexpr int C   = 16384 >> N2L; // columns per column tile
    static constexpr int G   = 16384 >> N1L; // rows per row group
    static constexpr int TW  = pass_tw_entries(LG);
    static_assert(C * Plan<N2L>::TA == kQueueThreads && G * Plan<N1L>::TA == kQueueThreads, "four-step tiles");
    static_assert(TW <= 2048, "pass twiddles in LDS");
    static constexpr int64_t N = (int64_t)1 << LG;
    static constexpr int     LS = (LG + 1) / 2;
};

// column pass: forward pass A (src -> intermediate, times the pass twiddle) or inverse pass 2
// (intermediate -> dst, conj and the 1/N scale at the store)
template<int LG, bool INV>
__device__ __noinline__ void queue_cols(const QueueArgs& q, const QItem it_) {
    using F              = QFour<LG>;
    constexpr int CT     = INV ? 0 : F::TW;
    float*        lds    = g_qlds;
    const QItem   it     = uni(it_);
    float*        tmp    = uni(q.tmp) + (int64_t)blockIdx.x * uni(q.tmp_floats);
    const float2* twl    = uni(q.ptw) + uni(q.ptw_off[LG]);
    ColArgs       ca{};
    ca.twA = uni(q.tw) + ((1 << F::N2L) - 2), ca.n1_lg = F::N1L, ca.ntiles = (1 << F::N1L) / F::C, ca.tw_lg = LG;
    ca.src_size = F::N, ca.src_lp = ca.dst_lp = LG; // split vectors and intermediate (pack = N)
    if constexpr (!INV) {
        ca.src = uni(q.src) + uni(uni(q.src_offs)[it.first]), ca.dst = tmp, ca.scale = 1.0f;
        ca.twl = twl, ca.twh = twl + (1 << F::LS), ca.tw_ls = F::LS;
    } else {
        ca.src = tmp, ca.dst = uni(q.dst) + uni(uni(q.dst_offs)[it.first]), ca.conj_out = 1;
        ca.scale = uni(q.normalized) ? (float)(1.0 / (double)F::N) : 1.0f;
    }
    using LC    = LdsCarve<float, 2 * (1 << F::N2L) * F::C, TwTable<F::N2L>::FLOATS, 2 * CT>;
    float* ltw  = lds + LC::TOFF;
    float* lptw = ltw + TwTable<F::N2L>::FLOATS;
    float* lre  = lds + LC::DOFF;
    float* lim  = lre + (1 << F::N2L) * F::C;
    TwTable<F::N2L>::fill(ltw, ca.twA, threadIdx.x, kQueueThreads);
    if constexpr (CT > 0)
        PassTw<CT>::fill(lptw, ca.twl, ca.twh, ca.tw_ls, ca.tw_lg, threadIdx.x, kQueueThreads);
    __syncthreads();
    for (int64_t tile = 0; tile < ca.ntiles; ++tile) {
        col_tile<F::N2L, F::C, CT>(ca, tile, lre, lim, ltw, lptw);
        __syncthreads(); // LDS is reused by the next tile
    }
}

// row pass: forward pass B (intermediate -> dst, strided store in natural order) or inverse pass
// 1 (src -> intermediate, strided load, times the pass twiddle)
template<int LG, bool INV>
__device__ __noinline__ void queue_rows(const QueueArgs& q, const QItem it_) {
    using F            = QFour<LG>;
    constexpr int RT   = INV ? F::TW : 0;
    using S            = RowShape<F::N1L, F::G>;
    float*        lds  = g_qlds;
    const QItem   it   = uni(it_);
    float*        tmp  = uni(q.tmp) + (int64_t)blockIdx.x * uni(q.tmp_floats);
    const float2* twl  = uni(q.ptw) + uni(q.ptw_off[LG]);
    RowArgs       ra{};
    ra.twA = uni(q.tw) + ((1 << F::N1L) - 2), ra.rows_lg = F::N2L, ra.nrows = 1 << F::N2L, ra.order = 0, ra.tw_lg = LG;
    ra.src_size = F::N, ra.src_lp = ra.dst_lp = LG, ra.scale = 1.0f;
    if constexpr (!INV) {
        ra.src = tmp, ra.dst = uni(q.dst) + uni(uni(q.dst_offs)[it.first]);
    } else {
        ra.src = uni(q.src) + uni(uni(q.src_offs)[it.first]), ra.dst = tmp, ra.conj_in = 1;
        ra.twl = twl, ra.twh = twl + (1 << F::LS), ra.tw_ls = F::LS;
    }
    using LC    = LdsCarve<float, S::DATA, TwTable<F::N1L>::FLOATS, 2 * RT>;
    float* ltw  = lds + LC::TOFF;
    float* lptw = ltw + TwTable<F::N1L>::FLOATS;
    float* lre  = lds + LC::DOFF;
    TwTable<F::N1L>::fill(ltw, ra.twA, threadIdx.x, kQueueThreads);
    if constexpr (RT > 0)
        PassTw<RT>::fill(lptw, ra.twl, ra.twh, ra.tw_ls, ra.tw_lg, threadIdx.x, kQueueThreads);
    __syncthreads();
    for (int64_t wg = 0; wg < ra.nrows / F::G; ++wg) {
        if constexpr (!INV)
            row_group<F::N1L, F::G, LD_CONTIG, ST_STRIDED, 0, true, 0>(ra, wg, lre, S::im(lre), ltw, lptw);
        else
            row_group<F::N1L, F::G, LD_STRIDED, ST_CONTIG, RT, true, 0>(ra, wg, lre, S::im(lre), ltw, lptw);
        __syncthreads(); // LDS is reused by the next group
    }
}

template<int LG>
__device__ __forceinline__ void queue_four(const QueueArgs& q, const QItem it) {
    if (!it.inv) {
        queue_cols<LG, false>(q, it);
        queue_rows<LG, false>(q, it);
    } else {
        queue_rows<LG, true>(q, it);
        queue_cols<LG, true>(q, it);
    }
}

__global__ __launch_bounds__(kQueueThreads) void k_queue(QueueArgs q) {
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0)
            s_item = (int)atomicAdd(q.counter, 1u);
        __syncthreads();
        const int it = s_item;
        if (it >= q.nitems)
            break; // uniform: every wavefront leaves together once the queue is drained
        const QItem w = q.items[it];
        if (w.kind == QK_SINGLE) {
            switch (w.la) {
            case 1: queue_single<1>(q, w); break;
            case 2: queue_single<2>(q, w); break;
            case 3: queue_single<3>(q, w); break;
            case 4: queue_single<4>(q, w); break;
            case 5: queue_single<5>(q, w); break;
            case 6: queue_single<6>(q, w); break;
            case 7: queue_single<7>(q, w); break;
            case 8: queue_single<8>(q, w); break;
            case 9: queue_single<9>(q, w); break;
            case 10: queue_single<10>(q, w); break;
            case 11: queue_single<11>(q, w); break;
            case 12: queue_single<12>(q, w); break;
            case 13: queue_single<13>(q, w); break;
            case 14: queue_single<14>(q, w); break;
            default: break;
            }
        } else {
            switch (w.la) {
            case 15: queue_four<15>(q, w); break;
            case 16: queue_four<16>(q, w); break;
            case 17: queue_four<17>(q, w); break;
            case 18: queue_four<18>(q, w); break;
            case 19: queue_four<19>(q, w); break;
            case 20: queue_four<20>(q, w); break;
            default: break;
            }
        }
        __syncthreads(); // LDS and s_item are reused by the next item
    }
}

__global__ void k_reset_counter(uint32_t* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *c = 0u;
}

} // namespace pcxk

namespace pcxd {

void launch_queue(const QueueArgs& a, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_queue, dim3(grid), dim3(kQueueThreads), 0, st, a);
}
void launch_reset_counter(uint32_t* c, hipStream_t st) { hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, c); }
void launch_vexpr(const VExprArgsT<float>& a, int64_t nwg, hipStream_t st) {
    hipLaunchKernelGGL(k_vexpr<float>, dim3((unsigned)nwg), dim3(256), 0, st, a);
}
void launch_vexpr(const VExprArgsT<double>& a, int64_t nwg, hipStream_t st) {
    hipLaunchKernelGGL(k_vexpr<double>, dim3((unsigned)nwg), dim3(256), 0, st, a);
}

} // namespace pcxd
