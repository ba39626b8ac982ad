// k_misc.hip -- the persistent uneven-batch queue kernel and the vector-expression kernel.
#include "dispatch.hpp"

namespace pcxk {

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
