// queue_kernels.hpp -- work items of the persistent queue kernel for batches of uneven transform
// sizes (the kernel, k_queue, is in k_misc.hip).
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

} // namespace pcxk
