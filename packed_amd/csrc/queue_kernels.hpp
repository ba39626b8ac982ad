// queue_kernels.hpp -- work items of the persistent queue kernel for batches of uneven transform
// sizes (the kernel, k_queue, is in k_misc.hip).
//
// One launch serves every transform of 2^1 .. 2^20 points.  Items are sorted largest first
// (longest-processing-time order); each 1024-thread workgroup takes the next item from a device
// counter (atomicAdd) until the counter passes the end, so every wavefront reaches the exit.
//  * single-pass items: up to 1024 / (2^la / 16) transforms of one size <= 2^14 and one direction,
//    run by the single-pass Stockham path of k_row;
//  * four-step items: one whole transform of 2^15 .. 2^20 points.  The workgroup runs all column
//    tiles, then all row tiles, through an intermediate of its own (2 * 2^20 floats per
//    workgroup), so no item waits for another workgroup.
// Transforms are addressed through per-transform offset tables in split re/im layout (pack = N).
#pragma once
#include "fft_kernels.hpp"

namespace pcxk {

constexpr int kQueueThreads   = 1024;
constexpr int kQueueMaxSingle = 14; // single-pass items: 2^1 .. 2^14
constexpr int kQueueMaxLog2   = 20; // the persistent launch serves 2^1 .. 2^20
constexpr int kQueueMinFour   = kQueueMaxSingle + 1;

enum : int32_t { QK_SINGLE = 0, QK_FOUR = 1 };

struct QItem {
    int32_t kind;  // QK_SINGLE or QK_FOUR
    int32_t la;    // log2 of the transform size
    int32_t inv;   // inverse (conj in / conj out)
    int32_t first; // index of the first transform in the offset tables
    int32_t count; // transforms in the item (1 for a four-step item)
};

struct QueueArgs {
    const float*   src;
    float*         dst;
    float*         tmp;        // four-step intermediates, tmp_floats per workgroup
    int64_t        tmp_floats;
    const int64_t* src_offs;   // per transform, in item order
    const int64_t* dst_offs;
    const QItem*   items;
    int32_t        nitems;
    int32_t        normalized;
    uint32_t*      counter;    // zeroed before every launch
    const float2*  tw;         // W_A tables for A = 2, 4, ..., 2^14 back to back (table A at A - 2)
    const float2*  ptw;        // pass tables W_N of N = 2^15 .. 2^20: lo (2^ls entries) then hi
    int32_t        ptw_off[kQueueMaxLog2 + 1]; // entry offset of N = 2^lg's tables in ptw
};

// The four-step split of 2^lg in the queue: rows of 2^q_n1(lg) points, columns of the rest;
// 1024-thread tiles of 16384 points (C = 16384 / N2 columns, G = 16384 / N1 rows).
constexpr int q_n1(int lg) { return lg == 15 ? 8 : (lg <= 19 ? 9 : 10); }
// Points per thread of a single-pass item: 16 (so every size uses all 1024 threads), or the
// whole transform below 16 points.
constexpr int q_ppt(int la) { return la >= 4 ? 16 : 0; }

} // namespace pcxk
