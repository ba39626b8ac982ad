// dispatch.hpp -- kernel launchers and the dispatch tables of the gfx950 FFT path.
//
// The tables (which kernel instantiation serves which size, order, precision and shape) are
// defined in the k_*.hip translation units, each of which instantiates its own kernels, so the
// library compiles in parallel.  pcx_gpu.hip (the C ABI) only reads them.
#pragma once
#include "fft_kernels.hpp"
#include "fxcorr_kernels.hpp"
#include "queue_kernels.hpp"
#include "vexpr_kernels.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <utility>

namespace pcxd {
using namespace pcxk;

constexpr int kMaxSinglePass   = 14; // N <= 16384 runs in one LDS-resident pass
constexpr int kMaxLog2         = 24; // four-step limit: N2 <= 2048 columns, N1 <= 16384
constexpr int kMaxSinglePass64 = 13; // FP64: N <= 8192 in one pass
constexpr int kMaxLog2_64      = 23; // FP64 four-step limit
constexpr int kDifLog2         = 15; // one-pass decimation-in-frequency size
constexpr int kMaxAcross       = 13; // transforms across at most 8192 vectors (float)
constexpr int kMaxAcross64     = 11; // and 2048 vectors (double)

template<class T>
using RowLaunchT = void (*)(const RowArgsT<T>&, int64_t nrows, int64_t grid_cap, hipStream_t);
template<class T>
using ColLaunchT = void (*)(const ColArgsT<T>&, int64_t grid_cap, hipStream_t);
using RowLaunch    = RowLaunchT<float>;
using ColLaunch    = ColLaunchT<float>;
using XcLaunch     = void (*)(const XcArgs&, hipStream_t);
using DifLaunch    = void (*)(const RowArgs&, int64_t batch, hipStream_t);
using AcrossLaunch   = void (*)(const AcrossArgs&, int64_t, hipStream_t);
using AcrossLaunch64 = void (*)(const AcrossArgsT<double>&, int64_t, hipStream_t);

// grid_cap > 0 bounds the grid; the kernels then loop over their work (grid-stride).  Grids are
// rounded up to a multiple of 8 for the XCD-aware item order (spare workgroups exit at once).
inline int64_t grid8(int64_t n) { return (n + 7) & ~(int64_t)7; }
// One dispatch covers at most 2^32 - 1 work-items (the 32-bit grid size of an AQL packet).
// Grid-stride kernels take a grid capped to that (a multiple of 8) and loop over the rest.
inline int64_t fit_grid(int64_t nwg, int64_t threads) {
    const int64_t cap = ((int64_t)0xFFFFFFFF / threads) & ~(int64_t)7;
    return nwg < cap ? nwg : cap;
}
inline int64_t pidx64(int64_t i, int lp) { return i + ((i >> lp) << lp); }

template<int LA, int G, int LD, int ST, int TW, bool UNI = (Plan<LA>::TA >= 64), int ORD = 0, int PREC = PREC_FP32,
         class T = float>
void launch_row(const RowArgsT<T>& a, int64_t nrows, int64_t grid_cap, hipStream_t st) {
    int64_t nwg = (nrows + G - 1) / G;
    if (grid_cap > 0 && nwg > grid_cap)
        nwg = grid_cap;
    nwg = fit_grid(grid8(nwg), G * Plan<LA, false, T>::TA);
    hipLaunchKernelGGL((k_row<LA, G, LD, ST, TW, UNI, ORD, PREC, T>), dim3((unsigned)nwg),
                       dim3(G * Plan<LA, false, T>::TA), 0, st, a);
}
template<int LA, int C, int TW, int PREC = PREC_FP32, class T = float, int PX = 0>
void launch_col(const ColArgsT<T>& a, int64_t grid_cap, hipStream_t st) {
    constexpr int NT  = C * Plan<LA, false, T, PX>::TA;
    int64_t       nwg = a.ntiles;
    if (grid_cap > 0 && nwg > grid_cap)
        nwg = grid_cap;
    nwg = fit_grid(grid8(nwg), NT);
    hipLaunchKernelGGL((k_col<LA, C, TW, PREC, T, PX>), dim3((unsigned)nwg), dim3(NT), 0, st, a);
}
template<int LA, int PREC = PREC_FP32>
void launch_xc(const XcArgs& a, hipStream_t st) {
    constexpr int G     = xc_rows<LA>();
    constexpr int NT    = G * xc_threads<LA>();
    const int64_t total = a.nblocks * a.channels;
    // 32-bit buffer offsets cover a channel when its packed footprint stays below 2^29 floats
    int  lp    = a.src_lp > a.dst_lp ? a.src_lp : a.dst_lp;
    bool small = pidx64(a.len + (int64_t)(1 << LA), lp) < ((int64_t)1 << 29) && a.src_lp < 29 && a.dst_lp < 29;
    // one workgroup per G blocks; launches of at most 2^32 - 1 work-items, in block order
    const int64_t per = fit_grid(grid8((total + G - 1) / G), NT) * G;
    XcArgs        s   = a;
    const int64_t end = a.blk1 > 0 ? a.blk1 : total; // a.blk0 .. a.blk1 - 1, or every block
    for (s.blk0 = a.blk0; s.blk0 < end; s.blk0 += per) {
        s.blk1            = std::min(end, s.blk0 + per);
        const int64_t nwg = grid8((s.blk1 - s.blk0 + G - 1) / G);
        if (a.order) // a fresh ticket counter and read flags for this launch (a.order holds nwg + 1 words)
            hipLaunchKernelGGL(k_xc_order_clear, dim3((unsigned)std::min<int64_t>((nwg + 256) / 256, 1024)), dim3(256),
                               0, st, a.order, nwg + 1);
        if (a.order && small)
            hipLaunchKernelGGL((k_fxcorr<LA, G, true, PREC, true>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
        else if (a.order)
            hipLaunchKernelGGL((k_fxcorr<LA, G, false, PREC, true>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
        else if (small)
            hipLaunchKernelGGL((k_fxcorr<LA, G, true, PREC>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
        else
            hipLaunchKernelGGL((k_fxcorr<LA, G, false, PREC>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
    }
}

// Single-pass variants: 0 natural in/out; 1, 2 permuted store (forward, bit_reversed /
// unordered); 3, 4 permuted load (inverse, bit_reversed / unordered).
using RowSet = std::array<RowLaunch, 5>;
inline int single_variant(int inverse, int order) { return order == 0 ? 0 : (inverse ? 2 : 0) + order; }

// Pass-twiddle table capacity in LDS for N = 2^LG (2: too large, read from global memory).
template<int LG>
constexpr int tw_cap() {
    return pass_tw_entries(LG) <= 2048 ? pass_tw_entries(LG) : 2;
}
template<int LG>
constexpr int tw_cap64() {
    return pass_tw_entries(LG) <= 512 ? pass_tw_entries(LG) : 2;
}

// Three-pass transform of 2^21 .. 2^24 points (FP64: 2^21 .. 2^23) in natural order:
// N = N1 * N2, N1 = Na * Nb with Nb = 2^7:
//   1. columns of length N2 (stride N1), times W_N^(n1*k2), into an intermediate with one
//      N1-point row (pack N1) per k2;
//   2. in each row, columns of length Nb (stride Na), times W_N1^(na*kb), in place;
//   3. rows of Na points, read in (kb, k2) order and stored to k2 + N2*kb + N2*Nb*ka.
// Every pass moves 128-byte or wider segments, where the natural-order four-step row pass of
// 2^11..2^14-point rows stores 8 to 2 rows (32 to 8 bytes) per segment.
constexpr int kThreeMinLog2 = 21;
constexpr int kThreeNbLog2  = 7;
template<class T>
struct ThreePassT {
    int            lg;
    int            n2_lg; // pass 1 column length
    int            na_lg; // pass 3 row length
    int            c1;    // pass 1 columns per tile
    int            c2;    // pass 2 columns per tile
    ColLaunchT<T>  col1;  // pass 1 (with the W_N pass twiddle)
    ColLaunchT<T>  col2;  // pass 2 (with the W_N1 pass twiddle)
    RowLaunchT<T>  row3;  // pass 3
};
template<class T>
struct ThreeList {
    const ThreePassT<T>* e;
    int                  count;
};
ThreeList<float>  three_pass_tab();   // 2^21 .. 2^24
ThreeList<double> three_pass64_tab(); // 2^21 .. 2^23
template<class T>
inline const ThreePassT<T>* three_of(int lg) {
    ThreeList<T> l;
    if constexpr (sizeof(T) == 8)
        l = three_pass64_tab();
    else
        l = three_pass_tab();
    for (int k = 0; k < l.count; ++k)
        if (l.e[k].lg == lg)
            return &l.e[k];
    return nullptr;
}

// Four-step shapes: N = 2^lg = N1 * N2 with N1 = 2^n1 (row length, G rows per workgroup,
// G * N1 / PPT <= 1024 threads) and N2 = 2^n2 (column length, C columns per tile, C * N2 <= 16384
// for LDS).  rows: forward with strided store (natural order) or permuted store (bit_reversed,
// unordered); inverse with strided or permuted load and the pass twiddle.  cols: {inverse,
// forward}.
template<class T>
struct FourStepT {
    int                          lg, n1, n2, g, c;
    std::array<RowLaunchT<T>, 6> rows;
    std::array<ColLaunchT<T>, 2> cols;
};
using FourStep = FourStepT<float>;

// Split-precision matrix-core variants of the default shapes for 2^15 .. 2^20 (natural order):
// index [precision - 1], rows {forward, inverse}, cols {inverse, forward}
struct FourStepMfma {
    int                                     lg;
    int                                     n1, n2; // the split its kernels are built for
    int                                     c;      // columns per tile of its column kernels
    std::array<std::array<RowLaunch, 2>, 2> rows;
    std::array<std::array<ColLaunch, 2>, 2> cols;
};

template<class T>
struct ShapeList {
    const FourStepT<T>* data;
    int                 count;
};

// ---- tables, defined in the k_*.hip translation units.  Each unit keeps its table in an
// anonymous namespace (a table visible outside it would also be emitted for the device, where
// its host launchers do not exist) and returns it from a host accessor.
// index LA-1 for LA in [1, 14]
const std::array<RowSet, kMaxSinglePass>& single_tab();
// split-precision single-pass variants (every order, both directions, as single_tab's sets):
// index [precision - 1][LA - 4]; natural order (k_mfma.hip), the permuted orders (k_mfma_p1.hip)
using MfmaSets = std::array<std::array<RowSet, kMaxSinglePass - 3>, 2>;
const std::array<std::array<RowLaunch, kMaxSinglePass - 3>, 2>& single_mfma_tab();
const MfmaSets&                                                 single_mfma_perm_tab();
const std::array<XcLaunch, kMaxSinglePass>& xc_tab();                         // index LA-1
const std::array<std::array<XcLaunch, kMaxSinglePass - 3>, 2>& xc_mfma_tab(); // [prec - 1][LA - 4]
const std::array<std::array<RowLaunchT<double>, 5>, kMaxSinglePass64>& single64_tab(); // index LA-1
// four-step shapes: the defaults (first match per size), then the alternatives of the tuning
// sweeps
ShapeList<float>  shapes_default();
ShapeList<float>  shapes_alt();
ShapeList<double> shapes64_default();
ShapeList<double> shapes64_alt();
const std::array<FourStepMfma, 6>& shapes_mfma();
// permuted-order row kernels of those shapes (k_mfma_p4.hip): [shape][precision - 1]
// {forward bit_reversed, forward unordered, inverse bit_reversed, inverse unordered}
using MfmaPermRows = std::array<std::array<std::array<RowLaunch, 4>, 2>, 6>;
const MfmaPermRows& shapes_mfma_perm();
// split-precision four-step shapes of 2^21 .. 2^24 (k_mfma_big.hip), the splits and tiles of the
// FP32 defaults: [precision - 1][LG - 21]
using MfmaBig = std::array<std::array<FourStep, 4>, 2>;
const MfmaBig& shapes_mfma_big();
// one-pass k_dif for N = 2^15, index [order]
const std::array<DifLaunch, 3>& dif_tab();
constexpr int dif_core_log2(int order) { return order == 0 ? 14 : 13; }
const std::array<AcrossLaunch, kMaxAcross>&     across_tab();   // index LA-1
const std::array<AcrossLaunch64, kMaxAcross64>& across64_tab(); // index LA-1

// ---- launchers of the remaining kernels (k_misc.hip)
void launch_queue(const QueueArgs& a, int grid, hipStream_t st);
void launch_reset_counter(uint32_t* c, hipStream_t st);
void launch_vexpr(const VExprArgsT<float>& a, int64_t nwg, hipStream_t st);
void launch_vexpr(const VExprArgsT<double>& a, int64_t nwg, hipStream_t st);

} // namespace pcxd
