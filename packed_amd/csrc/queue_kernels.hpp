// queue_kernels.hpp -- persistent work-queue kernel for batches of uneven transform sizes.
//
// One launch serves every single-pass transform of up to 4096 points.  Items are groups of
// transforms of one size and one direction, sorted largest first; each 256-thread workgroup takes
// the next item from a device counter (atomicAdd), runs it with the single-pass Stockham path of
// k_row and goes back to the queue until it is empty, so every wavefront reaches the exit.
// Transforms are addressed through per-transform offset tables in split re/im layout (pack = N).
#pragma once
#include "fft_kernels.hpp"

namespace pcxk {

constexpr int kQueueThreads = 256;
constexpr int kQueueMaxLog2 = 12;

struct QItem {
    int32_t la;    // log2 of the transform size
    int32_t inv;   // inverse (conj in / conj out)
    int32_t first; // index of the first transform in the offset tables
    int32_t count; // transforms in the item (<= kQueueThreads / (2^la / points per thread))
};

struct QueueArgs {
    const float*   src;
    float*         dst;
    const int64_t* src_offs; // per transform, in item order
    const int64_t* dst_offs;
    const QItem*   items;
    int32_t        nitems;
    int32_t        normalized;
    uint32_t*      counter; // zeroed before every launch
    const float2*  tw;      // W_A tables for A = 2, 4, ..., 4096 back to back (table A at A - 2)
};

template<int LA>
constexpr int queue_rows() {
    return kQueueThreads / Plan<LA>::TA;
}

template<int LA>
__device__ __forceinline__ void queue_item(const QueueArgs& q, const QItem& it, float* lds) {
    constexpr int G = queue_rows<LA>();
    RowArgs       p{};
    p.src      = q.src;
    p.dst      = q.dst;
    p.src_offs = q.src_offs + it.first;
    p.dst_offs = q.dst_offs + it.first;
    p.src_size = (int64_t)1 << LA;
    p.nrows    = it.count;
    p.src_lp = p.dst_lp = LA;
    p.rows_lg           = 0;
    p.order             = 0;
    p.conj_in = p.conj_out = it.inv;
    p.scale                = (it.inv && q.normalized) ? 1.0f / (float)(1 << LA) : 1.0f;
    p.twA                  = q.tw + ((1 << LA) - 2);
    float* ltw             = lds + RowShape<LA, G>::DATA;
    TwTable<LA>::fill(ltw, p.twA, threadIdx.x, kQueueThreads);
    __syncthreads();
    row_group<LA, G, LD_CONTIG, ST_CONTIG, 0, (Plan<LA>::TA >= 64), 0, PREC_FP32, float>(p, 0, lds, RowShape<LA, G>::im(lds),
                                                                                         ltw, nullptr);
}

template<int... L>
constexpr int queue_lds_floats(std::integer_sequence<int, L...>) {
    int m = 0;
    ((m = m > RowShape<L + 1, queue_rows<L + 1>()>::DATA + TwTable<L + 1>::FLOATS
              ? m
              : RowShape<L + 1, queue_rows<L + 1>()>::DATA + TwTable<L + 1>::FLOATS),
     ...);
    return m;
}
constexpr int kQueueLds = queue_lds_floats(std::make_integer_sequence<int, kQueueMaxLog2>{});

__global__ __launch_bounds__(kQueueThreads) void k_queue(QueueArgs q) {
    __shared__ float lds[kQueueLds];
    __shared__ int   s_item;
    for (;;) {
        if (threadIdx.x == 0)
            s_item = (int)atomicAdd(q.counter, 1u);
        __syncthreads();
        const int it = s_item;
        if (it >= q.nitems)
            break; // uniform: every wavefront leaves together once the queue is drained
        const QItem w = q.items[it];
        switch (w.la) {
        case 1: queue_item<1>(q, w, lds); break;
        case 2: queue_item<2>(q, w, lds); break;
        case 3: queue_item<3>(q, w, lds); break;
        case 4: queue_item<4>(q, w, lds); break;
        case 5: queue_item<5>(q, w, lds); break;
        case 6: queue_item<6>(q, w, lds); break;
        case 7: queue_item<7>(q, w, lds); break;
        case 8: queue_item<8>(q, w, lds); break;
        case 9: queue_item<9>(q, w, lds); break;
        case 10: queue_item<10>(q, w, lds); break;
        case 11: queue_item<11>(q, w, lds); break;
        case 12: queue_item<12>(q, w, lds); break;
        default: break;
        }
        __syncthreads(); // LDS and s_item are reused by the next item
    }
}

__global__ void k_reset_counter(uint32_t* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *c = 0u;
}

} // namespace pcxk
