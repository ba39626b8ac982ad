// pcx_gpu.hip -- C ABI of the gfx950 FFT / overlap-save cross-correlation path (include/pcx_gpu.h).
//
// Plans own device twiddle tables and, for N > 16384, a scratch buffer for the four-step
// intermediate.  Everything is launched asynchronously on the caller's stream.
#include "pcx_gpu.h"

#include "fft_kernels.hpp"
#include "fxcorr_kernels.hpp"
#include "vexpr_kernels.hpp"
#include "queue_kernels.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>
#include <algorithm>

using namespace pcxk;

namespace {

thread_local std::string g_err;

int fail(int code, std::string msg) {
    g_err = std::move(msg);
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return fail(PCX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));              \
    } while (0)

constexpr int kMaxSinglePass = 14; // N <= 16384 runs in one LDS-resident pass
constexpr int kMaxLog2       = 24; // four-step limit: N2 <= 2048 columns, N1 <= 16384
constexpr int kMaxSinglePass64 = 13; // FP64: N <= 8192 in one pass
constexpr int kMaxLog2_64      = 23; // FP64 four-step limit

bool    is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
int64_t pidx64(int64_t i, int lp) { return i + ((i >> lp) << lp); }
int  ilog2(int64_t v) {
    int l = 0;
    while (((int64_t)1 << l) < v)
        ++l;
    return l;
}

std::string s(int64_t v) { return std::to_string(v); }

// fft.hpp:1806-1812
int check_size(int64_t n) {
    if (n > 1 && is_pow2(n))
        return PCX_SUCCESS;
    return fail(PCX_ERR_INVALID_SIZE, "fft_size (which is  " + s(n) + ") is not an integer power of two");
}
// fft.hpp:1814-1826
int check_sub_size(int64_t sub) {
    if (sub >= 8 && is_pow2(sub))
        return PCX_SUCCESS;
    if (sub < 8)
        return fail(PCX_ERR_INVALID_SUB_SIZE,
                    "fft_sub_size (which is  " + s(sub) + ") is smaller than minimum (which is 8)");
    return fail(PCX_ERR_INVALID_SUB_SIZE, "fft_sub_size (which is  " + s(sub) + ") is not an integer power of two");
}
int check_pack(int64_t p, const char* what) {
    if (is_pow2(p) && p <= ((int64_t)1 << 40))
        return PCX_SUCCESS;
    return fail(PCX_ERR_INVALID_PACK, std::string(what) + " pack size (which is " + s(p) + ") is not a power of two");
}

template<class V>
hipError_t gen_twiddles(V* t, int64_t M, int64_t stride, int64_t count) {
    using T    = decltype(V{}.x);
    int64_t nb = (count + 255) / 256;
    hipLaunchKernelGGL(k_twiddles<T>, dim3((unsigned)nb), dim3(256), 0, 0, t, M, stride, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------- kernel dispatch tables
template<class T>
using RowLaunchT = void (*)(const RowArgsT<T>&, int64_t nrows, int64_t grid_cap, hipStream_t);
template<class T>
using ColLaunchT = void (*)(const ColArgsT<T>&, int64_t grid_cap, hipStream_t);
using RowLaunch  = RowLaunchT<float>;
using ColLaunch  = ColLaunchT<float>;
using XcLaunch  = void (*)(const XcArgs&, hipStream_t);

// grid_cap > 0 bounds the grid; the kernels then loop over their work (grid-stride).  Grids are
// rounded up to a multiple of 8 for the XCD-aware item order (spare workgroups exit at once).
int64_t grid8(int64_t n) { return (n + 7) & ~(int64_t)7; }
// One dispatch covers at most 2^32 - 1 work-items (the 32-bit grid size of an AQL packet).
// Grid-stride kernels take a grid capped to that (a multiple of 8) and loop over the rest.
int64_t fit_grid(int64_t nwg, int64_t threads) {
    const int64_t cap = ((int64_t)0xFFFFFFFF / threads) & ~(int64_t)7;
    return nwg < cap ? nwg : cap;
}
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
template<int LA, int C, int TW, int PREC = PREC_FP32, class T = float>
void launch_col(const ColArgsT<T>& a, int64_t grid_cap, hipStream_t st) {
    int64_t nwg = a.ntiles;
    if (grid_cap > 0 && nwg > grid_cap)
        nwg = grid_cap;
    nwg = fit_grid(grid8(nwg), C * Plan<LA, false, T>::TA);
    hipLaunchKernelGGL((k_col<LA, C, TW, PREC, T>), dim3((unsigned)nwg), dim3(C * Plan<LA, false, T>::TA), 0, st, a);
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
    for (s.blk0 = 0; s.blk0 < total; s.blk0 += per) {
        const int64_t nwg = grid8((std::min(per, total - s.blk0) + G - 1) / G);
        if (small)
            hipLaunchKernelGGL((k_fxcorr<LA, G, true, PREC>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
        else
            hipLaunchKernelGGL((k_fxcorr<LA, G, false, PREC>), dim3((unsigned)nwg), dim3(NT), 0, st, s);
    }
}

// Single-pass variants: 0 natural in/out; 1, 2 permuted store (forward, bit_reversed /
// unordered); 3, 4 permuted load (inverse, bit_reversed / unordered).
using RowSet = std::array<RowLaunch, 5>;
template<int LA>
constexpr RowSet single_set() {
    constexpr int  G   = default_rows<LA>();
    constexpr bool UNI = Plan<LA>::TA >= 64;
    return {&launch_row<LA, G, LD_CONTIG, ST_CONTIG, 0, UNI, 0>, &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_BITREV>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_UNORD>, &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_BITREV>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_UNORD>};
}
int single_variant(int inverse, int order) { return order == 0 ? 0 : (inverse ? 2 : 0) + order; }
template<int... L>
constexpr auto single_table(std::integer_sequence<int, L...>) {
    return std::array<RowSet, sizeof...(L)>{single_set<L + 1>()...};
}
template<int... L>
constexpr auto xc_table(std::integer_sequence<int, L...>) {
    return std::array<XcLaunch, sizeof...(L)>{&launch_xc<L + 1>...};
}
// index LA-1 for LA in [1, 14]
const auto kSingle = single_table(std::make_integer_sequence<int, kMaxSinglePass>{});
// split-precision matrix-core variants (natural order): index [precision - 1][LA - 4] for
// precision BF16X3 / FP16X2 and LA in [4, 14]
template<int PREC, int... L>
constexpr auto single_mfma_table(std::integer_sequence<int, L...>) {
    return std::array<RowLaunch, sizeof...(L)>{
        &launch_row<L + 4, default_rows<L + 4>(), LD_CONTIG, ST_CONTIG, 0, (Plan<L + 4>::TA >= 64), 0, PREC>...};
}
const std::array<std::array<RowLaunch, kMaxSinglePass - 3>, 2> kSingleMfma = {
    single_mfma_table<PREC_BF16X3>(std::make_integer_sequence<int, kMaxSinglePass - 3>{}),
    single_mfma_table<PREC_FP16X2>(std::make_integer_sequence<int, kMaxSinglePass - 3>{})};
const auto kXc = xc_table(std::make_integer_sequence<int, kMaxSinglePass>{});
template<int PREC, int... L>
constexpr auto xc_mfma_table(std::integer_sequence<int, L...>) {
    return std::array<XcLaunch, sizeof...(L)>{&launch_xc<L + 4, PREC>...};
}
const std::array<std::array<XcLaunch, kMaxSinglePass - 3>, 2> kXcMfma = {
    xc_mfma_table<PREC_BF16X3>(std::make_integer_sequence<int, kMaxSinglePass - 3>{}),
    xc_mfma_table<PREC_FP16X2>(std::make_integer_sequence<int, kMaxSinglePass - 3>{})};

using PassSet = std::array<RowLaunch, 6>;
// Pass-twiddle table capacity in LDS for N = 2^LG (2: too large, read from global memory).
template<int LG>
constexpr int tw_cap() {
    return pass_tw_entries(LG) <= 2048 ? pass_tw_entries(LG) : 2;
}
// ---- FP64 (fft_unit<double, Order>): natural and bit_reversed orders (the reference has no FP64
// unordered kernel: avx2_fft.hpp:114-217 is written for float only)
template<int LA>
constexpr std::array<RowLaunchT<double>, 5> single_set64() {
    using D             = double;
    constexpr int  G    = default_rows<LA, D>();
    constexpr bool UNI  = Plan<LA, false, D>::TA >= 64;
    return {&launch_row<LA, G, LD_CONTIG, ST_CONTIG, 0, UNI, 0, PREC_FP32, D>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_BITREV, PREC_FP32, D>, nullptr,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_BITREV, PREC_FP32, D>, nullptr};
}
template<int... L>
constexpr auto single_table64(std::integer_sequence<int, L...>) {
    return std::array<std::array<RowLaunchT<double>, 5>, sizeof...(L)>{single_set64<L + 1>()...};
}
const auto kSingle64 = single_table64(std::make_integer_sequence<int, kMaxSinglePass64>{}); // index LA-1

// Four-step row-pass variants: forward with strided store (natural order) or permuted store
// (bit_reversed, unordered); inverse with strided or permuted load and the pass twiddle.
// The G rows of a workgroup lie in one transform (G <= N2), so buffer addressing applies.
template<int LA, int G, int LG>
constexpr PassSet pass_set() {
    constexpr int TW = tw_cap<LG>();
    return {&launch_row<LA, G, LD_CONTIG, ST_STRIDED, 0, true, 0>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, true, ORD_BITREV>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, true, ORD_UNORD>,
            &launch_row<LA, G, LD_STRIDED, ST_CONTIG, TW, true, 0>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, TW, true, ORD_BITREV>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, TW, true, ORD_UNORD>};
}
template<int LA, int C, int LG>
constexpr std::array<ColLaunch, 2> col_set() {
    return {&launch_col<LA, C, 0>, &launch_col<LA, C, tw_cap<LG>()>};
}

// The four-step shapes this build instantiates: N = 2^lg = N1 * N2 with N1 = 2^n1 (row length,
// G rows per workgroup, G * N1 / PPT <= 1024 threads) and N2 = 2^n2 (column length, C columns
// per tile, C * N2 <= 16384 for LDS).
template<class T>
struct FourStepT {
    int                              lg, n1, n2, g, c;
    std::array<RowLaunchT<T>, 6>     rows;
    std::array<ColLaunchT<T>, 2>     cols;
};
using FourStep = FourStepT<float>;
template<int LG, int N1, int N2, int G, int C>
FourStep shape() {
    static_assert(N1 + N2 == LG, "shape");
    return {LG, N1, N2, G, C, pass_set<N1, G, LG>(), col_set<N2, C, LG>()};
}
const std::array<FourStep, 27> kShapes = {
    // defaults (first match per lg): measured best on MI355X for 2^16 is N1 = 512, N2 = 128 with
    // 128-column tiles (512-byte load and store segments in the column pass; 128 KiB of LDS, one
    // workgroup of 1024 threads per CU): 0.778 ms per 2 GiB against 0.855 ms with 64 columns;
    // 2^15 (four-step forward, 64 rows per row workgroup) 1.59 against 1.70 ms per 2 GiB
    shape<15, 8, 7, 64, 128>(), shape<16, 9, 7, 32, 128>(), shape<16, 8, 8, 32, 64>(), shape<17, 9, 8, 32, 64>(), shape<18, 9, 9, 32, 32>(),
    // 2^19: 512-point rows, 32 per workgroup (128-byte store segments), 1.91 against 2.03 ms per
    // 2 GiB with 1024-point rows (the alternative below)
    shape<19, 9, 10, 32, 16>(), shape<20, 10, 10, 16, 16>(), shape<21, 11, 10, 8, 16>(), shape<22, 12, 10, 4, 16>(),
    // 2^23, 2^24: more rows per workgroup than the LDS of two workgroups per CU allows (they
    // spill, but store 4 / 2 floats per segment instead of 2 / 1): 5.03 / 5.87 ms per 2 GiB
    // forward against 7.47 / 8.94
    shape<23, 12, 11, 4, 8>(), shape<24, 14, 10, 2, 16>(),
    // alternatives for N = 2^16 explored by the tuning sweeps
    shape<16, 8, 8, 16, 64>(), shape<16, 8, 8, 64, 64>(), shape<16, 8, 8, 32, 32>(), shape<16, 8, 8, 16, 32>(),
    shape<16, 10, 6, 16, 64>(), shape<16, 7, 9, 64, 32>(), shape<16, 6, 10, 64, 16>(),
    shape<16, 9, 7, 32, 32>(), shape<16, 10, 6, 16, 32>(), shape<16, 9, 7, 32, 64>(), shape<15, 8, 7, 32, 64>(), shape<15, 8, 7, 32, 128>(),
    shape<19, 10, 9, 16, 32>(),
    // alternatives for the largest sizes explored by the sweeps
    shape<23, 13, 10, 2, 16>(), shape<24, 14, 10, 1, 16>(), shape<24, 13, 11, 2, 8>(),
};
// FP64 shapes, 2^14 .. 2^23, 256-512 threads per workgroup.  C and G are the doubles per column
// segment and per row store segment.  The defaults are the measured best of the tuning sweeps on
// MI355X (2 GiB batches, forward): wider segments win over occupancy, e.g. 2^20 with C = G = 8
// and 129-133 KiB per workgroup runs in 2.43 ms against 3.23 ms for C = 8, G = 2.  Pass-twiddle
// tables above 512 entries are read from global memory.
template<int LG>
constexpr int tw_cap64() {
    return pass_tw_entries(LG) <= 512 ? pass_tw_entries(LG) : 2;
}
template<int LG, int N1, int N2, int G, int C>
FourStepT<double> shape64() {
    static_assert(N1 + N2 == LG && G <= (1 << N2), "shape");
    using D          = double;
    constexpr int TW = tw_cap64<LG>();
    return {LG, N1, N2, G, C,
            {&launch_row<N1, G, LD_CONTIG, ST_STRIDED, 0, true, 0, PREC_FP32, D>,
             &launch_row<N1, G, LD_CONTIG, ST_PERM, 0, true, ORD_BITREV, PREC_FP32, D>, nullptr,
             &launch_row<N1, G, LD_STRIDED, ST_CONTIG, TW, true, 0, PREC_FP32, D>,
             &launch_row<N1, G, LD_PERM, ST_CONTIG, TW, true, ORD_BITREV, PREC_FP32, D>, nullptr},
            {&launch_col<N2, C, 0, PREC_FP32, D>, &launch_col<N2, C, TW, PREC_FP32, D>}};
}
const std::array<FourStepT<double>, 29> kShapes64 = {
    // defaults (first match per lg); forward ms for a 2 GiB batch measured on MI355X
    // 2^14 .. 2^16: one workgroup per CU with 512-byte column segments (2^14, 2^15: 64 columns)
    // and 256-byte segments in both passes at 2^16 (32 columns, 32 rows): 1.63 / 1.58 / 1.62 ms
    // against 1.68 / 1.68 / 1.73 ms with the two-workgroups-per-CU tiles listed below
    shape64<14, 7, 7, 32, 64>(),  shape64<15, 8, 7, 32, 64>(), shape64<16, 8, 8, 32, 32>(),
    shape64<17, 8, 9, 32, 16>(),  // 1.84 ms
    shape64<18, 10, 8, 8, 32>(),  // 1.89
    shape64<19, 10, 9, 8, 16>(),  // 2.41
    shape64<20, 10, 10, 8, 8>(),  // 2.43
    shape64<21, 11, 10, 4, 8>(),  // 3.11
    shape64<22, 12, 10, 4, 8>(),  // 3.12
    shape64<23, 13, 10, 2, 4>(),  // 4.91
    // alternatives explored by the sweeps
    shape64<17, 9, 8, 8, 16>(),   // 2.08
    shape64<17, 9, 8, 16, 16>(),  // 1.90 (inverse 2.05)
    shape64<18, 10, 8, 4, 16>(),  // 2.71
    shape64<18, 9, 9, 8, 8>(),    // 2.22
    shape64<18, 8, 10, 16, 4>(),  // 2.15
    shape64<19, 10, 9, 4, 8>(),   // 3.25
    shape64<19, 9, 10, 8, 4>(),   // 2.81
    shape64<20, 11, 9, 2, 8>(),   // 3.23
    shape64<20, 10, 10, 4, 4>(),  // 3.15
    shape64<20, 11, 9, 4, 8>(),   // 2.63
    shape64<21, 12, 9, 2, 8>(),   // 3.59
    shape64<22, 13, 9, 1, 8>(),   // 4.49
    shape64<23, 13, 10, 1, 4>(),  // 5.91
    shape64<14, 7, 7, 32, 32>(), shape64<15, 8, 7, 16, 32>(), shape64<16, 8, 8, 16, 16>(),
    shape64<17, 8, 9, 16, 8>(),   // 1.88
    shape64<18, 10, 8, 8, 16>(),  // 2.02
    shape64<19, 10, 9, 8, 8>()};  // 2.44
const FourStepT<double>* find_shape64(int lg, int n2, int g, int c) {
    for (const auto& f : kShapes64)
        if (f.lg == lg && (n2 == 0 || f.n2 == n2) && (g == 0 || f.g == g) && (c == 0 || f.c == c))
            return &f;
    return nullptr;
}

// split-precision matrix-core variants of the default shapes for 2^15 .. 2^20 (natural order):
// index [precision - 1], rows {forward, inverse}, cols {inverse, forward}
struct FourStepMfma {
    int                                     lg;
    int                                     n1, n2; // the split its kernels are built for
    int                                     c;      // columns per tile of its column kernels
    std::array<std::array<RowLaunch, 2>, 2> rows;
    std::array<std::array<ColLaunch, 2>, 2> cols;
};
template<int LG, int N1, int N2, int G, int C>
FourStepMfma shape_mfma() {
    static_assert(N1 + N2 == LG, "shape");
    return {LG, N1, N2, C,
            {{{&launch_row<N1, G, LD_CONTIG, ST_STRIDED, 0, true, 0, PREC_BF16X3>,
               &launch_row<N1, G, LD_STRIDED, ST_CONTIG, tw_cap<LG>(), true, 0, PREC_BF16X3>},
              {&launch_row<N1, G, LD_CONTIG, ST_STRIDED, 0, true, 0, PREC_FP16X2>,
               &launch_row<N1, G, LD_STRIDED, ST_CONTIG, tw_cap<LG>(), true, 0, PREC_FP16X2>}}},
            {{{&launch_col<N2, C, 0, PREC_BF16X3>, &launch_col<N2, C, tw_cap<LG>(), PREC_BF16X3>},
              {&launch_col<N2, C, 0, PREC_FP16X2>, &launch_col<N2, C, tw_cap<LG>(), PREC_FP16X2>}}}};
}
const std::array<FourStepMfma, 6> kShapesMfma = {shape_mfma<15, 8, 7, 32, 64>(),  shape_mfma<16, 9, 7, 32, 64>(),
                                                 shape_mfma<17, 9, 8, 32, 64>(),  shape_mfma<18, 9, 9, 32, 32>(),
                                                 shape_mfma<19, 9, 10, 32, 16>(), shape_mfma<20, 10, 10, 16, 16>()};
// One-pass decimation-in-frequency kernel (k_dif) for N = 2^15 = M * R: natural order with
// M = 2^14, R = 2; bit_reversed / unordered with M = 2^13, R = 4 (two workgroups per CU; the
// stores are contiguous there).  Measured on MI355X against the four-step for a 2 GiB batch:
// 1.55 vs 1.72 ms (natural), 1.50 vs 1.70 ms (bit_reversed).  For 2^16 and 2^17 the R-fold
// rereads cost more than the second pass saves (2.27 vs 1.71 ms at 2^16), so those sizes stay
// four-step.  Index [order].
using DifLaunch = void (*)(const RowArgs&, int64_t batch, hipStream_t);
template<int LM, int LR, int ORD>
void launch_dif(const RowArgs& a, int64_t batch, hipStream_t st) {
    int64_t nwg = ((batch + 7) / 8) * 8 * ((int64_t)1 << LR);
    hipLaunchKernelGGL((k_dif<LM, LR, ORD>), dim3((unsigned)nwg), dim3(Plan<LM, false, float, dif_ppt<LM>()>::TA), 0, st,
                       a);
}
constexpr int kDifLog2 = 15;
const std::array<DifLaunch, 3> kDif = {&launch_dif<14, 1, ORD_NORMAL>, &launch_dif<13, 2, ORD_BITREV>,
                                       &launch_dif<13, 2, ORD_UNORD>};
constexpr int dif_core_log2(int order) { return order == 0 ? 14 : 13; }

const FourStep* find_shape(int lg, int n2, int g, int c);
// The split kernels for the plan's shape: same size and the same N1 x N2 split (the plan's
// scratch layout and row / column counts follow its shape); only the default shapes have them.
const FourStepMfma* find_mfma(int lg, const FourStep* sh) {
    for (const auto& f : kShapesMfma)
        if (f.lg == lg && sh == find_shape(lg, 0, 0, 0) && f.n1 == sh->n1 && f.n2 == sh->n2)
            return &f;
    return nullptr;
}

const FourStep* find_shape(int lg, int n2, int g, int c) {
    for (const auto& f : kShapes)
        if (f.lg == lg && (n2 == 0 || f.n2 == n2) && (g == 0 || f.g == g) && (c == 0 || f.c == c))
            return &f;
    return nullptr;
}

} // namespace

// ---------------------------------------------------------------- plan
struct pcx_plan_s {
    int64_t n     = 0;
    int     lg    = 0;
    int     order = 0;
    bool    two_pass = false;
    int     n1_lg = 0, n2_lg = 0; // N = N1 * N2, N1 contiguous (row) length, N2 column length
    float2* tw_row = nullptr;     // W_N (single pass) or W_N1
    float2* tw_col = nullptr;     // W_N2
    float2* twl    = nullptr;     // W_N^e, e < 2^tw_ls
    float2* twh    = nullptr;     // W_N^(e << tw_ls)
    int     tw_ls  = 0;
    float*  scratch      = nullptr;
    int64_t scratch_cap  = 0; // transforms the scratch holds
    int64_t chunk        = 0; // transforms per four-step chunk
    const FourStep* shape = nullptr; // four-step kernels and tile sizes
    int     precision    = PCX_PREC_FP32;
    int64_t grid_col     = 2048; // grid caps of the two passes (persistent loop; 0: one workgroup per tile)
    int64_t grid_row     = 2048;
    int     nt_flags     = 0; // bit 0/1: streaming load/store of src/dst, bit 2/3: of the scratch
    float2* tw_dif       = nullptr; // W_M for the one-pass k_dif kernel
    bool    use_dif      = false;   // out-of-place transforms of 2^15 run as one k_dif pass
    // FP64 plans (pcx_plan_create_f64): the same roles in double precision
    bool     f64      = false;
    double2* tw64_row = nullptr;
    double2* tw64_col = nullptr;
    double2* twl64    = nullptr;
    double2* twh64    = nullptr;
    const FourStepT<double>* shape64 = nullptr;
    std::mutex mu;
};

extern "C" int pcx_version(void) { return 1; }
extern "C" const char* pcx_last_error(void) { return g_err.c_str(); }

extern "C" int pcx_plan_create(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan output pointer is null");
    *out = nullptr;
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_sub_size(sub_size))
        return e;
    if (order < 0 || order > 2)
        return fail(PCX_ERR_INVALID_ARGUMENT, "order (which is " + s(order) + ") is not a fft_order");
    int lg = ilog2(fft_size);
    if (lg > kMaxLog2)
        return fail(PCX_ERR_UNSUPPORTED, "fft_size (which is " + s(fft_size) + ") exceeds the supported maximum 2^24");
    if (order == PCX_ORDER_UNORDERED && lg < 5)
        return fail(PCX_ERR_UNSUPPORTED,
                    "unordered fft_size (which is " + s(fft_size) + ") is below the 32-point finisher size");
    auto* p  = new pcx_plan_s;
    p->n     = fft_size;
    p->lg    = lg;
    p->order = order;
    hipError_t he = hipSuccess;
    if (lg <= kMaxSinglePass) {
        he = hipMalloc(&p->tw_row, sizeof(float2) * fft_size);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw_row, fft_size, 1, fft_size);
    } else {
        p->two_pass = true;
        // shape: N2 = 2^min(lg/2, 10) columns-length by default; PCX_N2_LOG2 and PCX_TILE ("C,G")
        // select one of the other instantiated shapes (tuning sweeps)
        int n2s = 0, c = 0, g = 0; // 0: the default (first listed) shape for this size
        if (const char* env = std::getenv("PCX_N2_LOG2"))
            n2s = std::atoi(env);
        if (const char* env = std::getenv("PCX_TILE"))
            if (std::sscanf(env, "%d,%d", &c, &g) != 2)
                c = g = 0;
        p->shape = find_shape(lg, n2s, g, c);
        if (!p->shape)
            p->shape = find_shape(lg, n2s, 0, 0);
        if (!p->shape)
            p->shape = find_shape(lg, 0, 0, 0);
        if (!p->shape) {
            delete p;
            return fail(PCX_ERR_UNSUPPORTED, "no four-step kernels for fft_size " + s(fft_size));
        }
        p->n2_lg = p->shape->n2;
        p->n1_lg = p->shape->n1;
        p->tw_ls = (lg + 1) / 2;
        int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
        int64_t nl = (int64_t)1 << p->tw_ls, nh = fft_size >> p->tw_ls;
        he = hipMalloc(&p->tw_row, sizeof(float2) * n1);
        if (he == hipSuccess)
            he = hipMalloc(&p->tw_col, sizeof(float2) * n2);
        if (he == hipSuccess)
            he = hipMalloc(&p->twl, sizeof(float2) * nl);
        if (he == hipSuccess)
            he = hipMalloc(&p->twh, sizeof(float2) * nh);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw_row, n1, 1, n1);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw_col, n2, 1, n2);
        if (he == hipSuccess)
            he = gen_twiddles(p->twl, fft_size, 1, nl);
        if (he == hipSuccess)
            he = gen_twiddles(p->twh, fft_size, nl, nh);
        // transforms per four-step chunk: the scratch holds one chunk's intermediate
        int64_t chunk_bytes = (int64_t)2048 << 20;
        if (const char* env = std::getenv("PCX_CHUNK_MB"))
            chunk_bytes = (int64_t)std::atoll(env) << 20;
        p->chunk = chunk_bytes / (fft_size * 8);
        if (p->chunk < 1)
            p->chunk = 1;
        if (const char* env = std::getenv("PCX_GRID")) { // "col,row" grid caps for tuning sweeps
            long long ga = 0, gb = 0;
            if (std::sscanf(env, "%lld,%lld", &ga, &gb) == 2)
                p->grid_col = ga, p->grid_row = gb;
        }
        if (const char* env = std::getenv("PCX_NT"))
            p->nt_flags = std::atoi(env);
        if (lg == kDifLog2) {
            p->use_dif = true;
            if (const char* env = std::getenv("PCX_DIF"))
                p->use_dif = std::atoi(env) != 0;
            const int64_t m = (int64_t)1 << dif_core_log2(order);
            if (he == hipSuccess)
                he = hipMalloc(&p->tw_dif, sizeof(float2) * m);
            if (he == hipSuccess)
                he = gen_twiddles(p->tw_dif, m, 1, m);
        }
    }
    if (he == hipSuccess)
        he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        pcx_plan_destroy(p);
        return fail(PCX_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(he));
    }
    *out = p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_destroy(pcx_plan_t p) {
    if (!p)
        return PCX_SUCCESS;
    (void)hipFree(p->tw_row);
    (void)hipFree(p->tw_col);
    (void)hipFree(p->twl);
    (void)hipFree(p->twh);
    (void)hipFree(p->scratch);
    (void)hipFree(p->tw_dif);
    (void)hipFree(p->tw64_row);
    (void)hipFree(p->tw64_col);
    (void)hipFree(p->twl64);
    (void)hipFree(p->twh64);
    delete p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_create_f64(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan output pointer is null");
    *out = nullptr;
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_sub_size(sub_size))
        return e;
    if (order < 0 || order > 2)
        return fail(PCX_ERR_INVALID_ARGUMENT, "order (which is " + s(order) + ") is not a fft_order");
    if (order == PCX_ORDER_UNORDERED)
        return fail(PCX_ERR_UNSUPPORTED, "the unordered order is defined for float only (the reference's "
                                         "unsorted kernels are float-only)");
    int lg = ilog2(fft_size);
    if (lg > kMaxLog2_64)
        return fail(PCX_ERR_UNSUPPORTED,
                    "double fft_size (which is " + s(fft_size) + ") exceeds the supported maximum 2^23");
    auto* p  = new pcx_plan_s;
    p->n     = fft_size;
    p->lg    = lg;
    p->order = order;
    p->f64   = true;
    hipError_t he = hipSuccess;
    if (lg <= kMaxSinglePass64) {
        he = hipMalloc(&p->tw64_row, sizeof(double2) * fft_size);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_row, fft_size, 1, fft_size);
    } else {
        p->two_pass = true;
        // PCX_N2_LOG2 and PCX_TILE ("C,G") select another instantiated shape (tuning sweeps)
        int n2s = 0, c = 0, g = 0;
        if (const char* env = std::getenv("PCX_N2_LOG2"))
            n2s = std::atoi(env);
        if (const char* env = std::getenv("PCX_TILE"))
            if (std::sscanf(env, "%d,%d", &c, &g) != 2)
                c = g = 0;
        p->shape64 = find_shape64(lg, n2s, g, c);
        if (!p->shape64)
            p->shape64 = find_shape64(lg, 0, 0, 0);
        p->n1_lg = p->shape64->n1;
        p->n2_lg = p->shape64->n2;
        p->tw_ls = (lg + 1) / 2;
        int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
        int64_t nl = (int64_t)1 << p->tw_ls, nh = fft_size >> p->tw_ls;
        he = hipMalloc(&p->tw64_row, sizeof(double2) * n1);
        if (he == hipSuccess)
            he = hipMalloc(&p->tw64_col, sizeof(double2) * n2);
        if (he == hipSuccess)
            he = hipMalloc(&p->twl64, sizeof(double2) * nl);
        if (he == hipSuccess)
            he = hipMalloc(&p->twh64, sizeof(double2) * nh);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_row, n1, 1, n1);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_col, n2, 1, n2);
        if (he == hipSuccess)
            he = gen_twiddles(p->twl64, fft_size, 1, nl);
        if (he == hipSuccess)
            he = gen_twiddles(p->twh64, fft_size, nl, nh);
        int64_t chunk_bytes = (int64_t)2048 << 20;
        if (const char* env = std::getenv("PCX_CHUNK_MB"))
            chunk_bytes = (int64_t)std::atoll(env) << 20;
        p->chunk = chunk_bytes / (fft_size * 16);
        if (p->chunk < 1)
            p->chunk = 1;
    }
    if (he == hipSuccess)
        he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        pcx_plan_destroy(p);
        return fail(PCX_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(he));
    }
    *out = p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_element_size(pcx_plan_t p) { return p ? (p->f64 ? 8 : 4) : 0; }

extern "C" int64_t pcx_plan_size(pcx_plan_t p) { return p ? p->n : 0; }

extern "C" int pcx_plan_set_precision(pcx_plan_t p, int precision) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64)
        return fail(PCX_ERR_UNSUPPORTED, "a double plan has no precision modes");
    if (precision == PCX_PREC_FP32) {
        p->precision = precision;
        return PCX_SUCCESS;
    }
    if (precision != PCX_PREC_BF16X3 && precision != PCX_PREC_FP16X2)
        return fail(PCX_ERR_UNSUPPORTED, "precision " + s(precision) + " is not implemented");
    if (p->order != PCX_ORDER_NORMAL)
        return fail(PCX_ERR_UNSUPPORTED, "split-precision transforms are built for the normal order only");
    if (p->lg < 4 || p->lg > 20 || (p->two_pass && !find_mfma(p->lg, p->shape)))
        return fail(PCX_ERR_UNSUPPORTED, "split-precision transforms cover 2^4 .. 2^20 (default four-step shapes)");
    p->precision = precision;
    return PCX_SUCCESS;
}
extern "C" int pcx_plan_precision(pcx_plan_t p) { return p ? p->precision : -1; }
extern "C" int     pcx_plan_order(pcx_plan_t p) { return p ? p->order : -1; }

namespace {

// Switches a four-step plan to another instantiated shape, rebuilding the W_N1 / W_N2 tables
// when the split changes (the pass tables W_N do not depend on the split).
int plan_apply_shape(pcx_plan_t p, const FourStep* sh) {
    if (!p->two_pass || !sh || sh->lg != p->lg)
        return fail(PCX_ERR_INVALID_ARGUMENT, "shape does not match the plan");
    if (sh->n1 != p->n1_lg || sh->n2 != p->n2_lg) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(p->tw_row));
        HIPCHK(hipFree(p->tw_col));
        p->tw_row = p->tw_col = nullptr;
        int64_t n1 = (int64_t)1 << sh->n1, n2 = (int64_t)1 << sh->n2;
        HIPCHK(hipMalloc(&p->tw_row, sizeof(float2) * n1));
        HIPCHK(hipMalloc(&p->tw_col, sizeof(float2) * n2));
        HIPCHK(gen_twiddles(p->tw_row, n1, 1, n1));
        HIPCHK(gen_twiddles(p->tw_col, n2, 1, n2));
        HIPCHK(hipDeviceSynchronize());
    }
    p->shape = sh, p->n1_lg = sh->n1, p->n2_lg = sh->n2;
    return PCX_SUCCESS;
}

} // namespace

namespace {

// src_offs / dst_offs (optional, device arrays of `batch` element offsets) replace b * dist.
// T is the element type: float, or double for a plan made by pcx_plan_create_f64.
template<class T>
int run_batch(pcx_plan_t p, int inverse, int normalized, T* dst, int64_t dst_pack, int64_t dst_dist,
              const T* src, int64_t src_pack, int64_t src_dist, int64_t src_size, int64_t batch,
              hipStream_t st, const int64_t* src_offs = nullptr, const int64_t* dst_offs = nullptr) {
    constexpr bool F64 = std::is_same_v<T, double>;
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64 != F64)
        return fail(PCX_ERR_INVALID_ARGUMENT, p->f64 ? "the plan is for double data (use the _f64 functions)"
                                                     : "the plan is for float data (made by pcx_plan_create)");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (batch < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "batch (which is " + s(batch) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    const int64_t n = p->n;
    if (src_size > n)
        return fail(PCX_ERR_SIZE_MISMATCH,
                    "source size (which is " + s(src_size) + ") is bigger than fft size (which is " + s(n) + ")");
    if (src_size < 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "source size (which is " + s(src_size) + ") is negative");
    // spectra in unordered order keep pack max(P, 8) (avx2_fft.hpp:209-213, 227)
    if (p->order == PCX_ORDER_UNORDERED) {
        if (inverse && src_pack < 8)
            src_pack = 8;
        if (!inverse && dst_pack < 8)
            dst_pack = 8;
    }
    if ((const T*)dst == src && dst_pack != src_pack && (dst_pack > 8 || src_pack > 8))
        return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                s(src_pack) + ") to destination pack size (which is " + s(dst_pack) + ")");
    if (batch == 0)
        return PCX_SUCCESS;
    const int   src_lp = ilog2(src_pack), dst_lp = ilog2(dst_pack);
    const T     scale  = (inverse && normalized) ? (T)(1.0 / (double)n) : T(1);
    if (inverse)
        src_size = n;
    const vec2<T> *tw_row, *tw_col, *twl, *twh;
    if constexpr (F64)
        tw_row = p->tw64_row, tw_col = p->tw64_col, twl = p->twl64, twh = p->twh64;
    else
        tw_row = p->tw_row, tw_col = p->tw_col, twl = p->twl, twh = p->twh;

    if (!p->two_pass) {
        RowArgsT<T> a{};
        a.src = src, a.dst = dst, a.src_dist = src_dist, a.dst_dist = dst_dist;
        a.src_size = src_size, a.nrows = batch, a.src_lp = src_lp, a.dst_lp = dst_lp, a.rows_lg = 0;
        a.order = p->order, a.conj_in = inverse, a.conj_out = inverse, a.scale = scale;
        a.twA = tw_row, a.twl = nullptr, a.twh = nullptr, a.tw_ls = 0;
        a.src_offs = src_offs, a.dst_offs = dst_offs;
        if constexpr (F64)
            kSingle64[p->lg - 1][single_variant(inverse, p->order)](a, batch, 0, st);
        else if (p->precision != PCX_PREC_FP32)
            kSingleMfma[p->precision - 1][p->lg - 4](a, batch, 0, st);
        else
            kSingle[p->lg - 1][single_variant(inverse, p->order)](a, batch, 0, st);
        HIPCHK(hipGetLastError());
        return PCX_SUCCESS;
    }

    // ---- one pass, radix-R decimation in frequency + M-point Stockham (no intermediate)
    if constexpr (!F64)
    if (p->use_dif && p->precision == PCX_PREC_FP32 && (const float*)dst != src &&
        (!inverse || p->order == PCX_ORDER_NORMAL)) {
        RowArgs a{};
        a.src = src, a.dst = dst, a.src_dist = src_dist, a.dst_dist = dst_dist;
        a.src_size = src_size, a.nrows = batch, a.src_lp = src_lp, a.dst_lp = dst_lp;
        a.conj_in = inverse, a.conj_out = inverse, a.scale = scale;
        a.twA = p->tw_dif, a.twl = p->twl, a.twh = p->twh, a.tw_ls = p->tw_ls, a.tw_lg = p->lg;
        a.nt_store = (p->nt_flags >> 1) & 1;
        a.src_offs = src_offs, a.dst_offs = dst_offs;
        kDif[p->order](a, batch, st);
        HIPCHK(hipGetLastError());
        return PCX_SUCCESS;
    }

    // ---- four-step: N = N1 * N2 through a plan-owned scratch, `chunk` transforms at a time
    std::lock_guard<std::mutex> lock(p->mu);
    int64_t chunk = batch < p->chunk ? batch : p->chunk;
    if (p->scratch_cap < chunk) {
        if (p->scratch) {
            HIPCHK(hipStreamSynchronize(st));
            HIPCHK(hipFree(p->scratch));
            p->scratch = nullptr;
        }
        hipError_t he = hipMalloc(&p->scratch, sizeof(T) * 2 * n * chunk);
        if (he != hipSuccess)
            return fail(PCX_ERR_OUT_OF_MEMORY, "scratch allocation of " + s(2 * sizeof(T) * n * chunk) + " bytes failed");
        p->scratch_cap = chunk;
    }
    T* const scratch = reinterpret_cast<T*>(p->scratch);
    const FourStepT<T>* shape;
    if constexpr (F64)
        shape = p->shape64;
    else
        shape = p->shape;
    int           C    = shape->c;
    ColLaunchT<T> colf = shape->cols[inverse ? 0 : 1];
    RowLaunchT<T> rowf = shape->rows[(inverse ? 3 : 0) + p->order];
    if constexpr (!F64)
    if (p->precision != PCX_PREC_FP32) {
        const FourStepMfma* fm = find_mfma(p->lg, p->shape);
        if (!fm)
            return fail(PCX_ERR_UNSUPPORTED, "no split-precision kernels for this four-step shape");
        colf = fm->cols[p->precision - 1][inverse ? 0 : 1];
        rowf = fm->rows[p->precision - 1][inverse ? 1 : 0];
        C    = fm->c; // the split kernels keep their own column tile
    }
    const int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
    for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
        int64_t      cb = batch - b0 < chunk ? batch - b0 : chunk;
        const T*       cs  = src_offs ? src : src + b0 * src_dist;
        T*             cd  = dst_offs ? dst : dst + b0 * dst_dist;
        const int64_t* cso = src_offs ? src_offs + b0 : nullptr;
        const int64_t* cdo = dst_offs ? dst_offs + b0 : nullptr;
        ColArgsT<T>  ca{};
        RowArgsT<T>  ra{};
        ca.twA = tw_col, ca.n1_lg = p->n1_lg, ca.ntiles = cb * (n1 / C);
        ra.twA = tw_row, ra.rows_lg = p->n2_lg, ra.nrows = cb * n2, ra.order = p->order;
        ca.tw_lg = p->lg, ra.tw_lg = p->lg;
        if (!inverse) {
            // pass A: columns of src -> scratch, times W_N^(n1*k2)
            ca.src = cs, ca.src_dist = src_dist, ca.src_lp = src_lp, ca.src_size = src_size, ca.src_offs = cso;
            ca.dst = scratch, ca.dst_dist = 2 * n, ca.dst_lp = p->lg;
            ca.conj_in = 0, ca.conj_out = 0, ca.scale = T(1);
            ca.twl = twl, ca.twh = twh, ca.tw_ls = p->tw_ls;
            ca.nt_load = p->nt_flags & 1, ca.nt_store = (p->nt_flags >> 2) & 1;
            colf(ca, p->grid_col, st);
            // pass B: rows of scratch -> dst in the plan's order
            ra.src = scratch, ra.src_dist = 2 * n, ra.src_lp = p->lg, ra.src_size = n;
            ra.dst = cd, ra.dst_dist = dst_dist, ra.dst_lp = dst_lp, ra.dst_offs = cdo;
            ra.conj_in = 0, ra.conj_out = 0, ra.scale = T(1);
            ra.twl = nullptr, ra.twh = nullptr;
            ra.nt_load = (p->nt_flags >> 3) & 1, ra.nt_store = (p->nt_flags >> 1) & 1;
            rowf(ra, ra.nrows, p->grid_row, st);
        } else {
            // pass 1: rows (over k1) of the spectrum -> scratch, times W_N^(n1*k2)
            ra.src = cs, ra.src_dist = src_dist, ra.src_lp = src_lp, ra.src_size = n, ra.src_offs = cso;
            ra.dst = scratch, ra.dst_dist = 2 * n, ra.dst_lp = p->lg;
            ra.conj_in = 1, ra.conj_out = 0, ra.scale = T(1);
            ra.twl = twl, ra.twh = twh, ra.tw_ls = p->tw_ls;
            ra.nt_load = p->nt_flags & 1, ra.nt_store = (p->nt_flags >> 2) & 1;
            rowf(ra, ra.nrows, p->grid_row, st);
            // pass 2: columns (over k2) of scratch -> dst, natural order
            ca.src = scratch, ca.src_dist = 2 * n, ca.src_lp = p->lg, ca.src_size = n;
            ca.dst = cd, ca.dst_dist = dst_dist, ca.dst_lp = dst_lp, ca.dst_offs = cdo;
            ca.conj_in = 0, ca.conj_out = 1, ca.scale = scale;
            ca.twl = nullptr, ca.twh = nullptr;
            ca.nt_load = (p->nt_flags >> 3) & 1, ca.nt_store = (p->nt_flags >> 1) & 1;
            colf(ca, p->grid_col, st);
        }
        HIPCHK(hipGetLastError());
    }
    return PCX_SUCCESS;
}

int64_t footprint(int64_t n, int64_t pack) { return ((n + pack - 1) / pack) * pack * 2; }

} // namespace

extern "C" int pcx_fft(pcx_plan_t p, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
                       int64_t src_size, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 0, 0, dst, dst_pack, 0, src, src_pack, 0, src_size, 1, (hipStream_t)stream);
}

extern "C" int pcx_ifft(pcx_plan_t p, int normalized, float* dst, int64_t dst_pack, const float* src,
                        int64_t src_pack, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 1, normalized, dst, dst_pack, 0, src, src_pack, 0, p->n, 1, (hipStream_t)stream);
}

namespace {
template<class T>
int fft_batch(pcx_plan_t p, int inverse, int normalized, T* dst, int64_t dst_pack, int64_t dst_dist, const T* src,
              int64_t src_pack, int64_t src_dist, int64_t src_size, int64_t batch, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (batch > 1) {
        int64_t dp = dst_pack, sp = src_pack;
        if (p->order == PCX_ORDER_UNORDERED) {
            if (!inverse && dp < 8)
                dp = 8;
            if (inverse && sp < 8)
                sp = 8;
        }
        // a forward source holds src_size values (zero-extended to the fft size)
        const int64_t src_n = inverse ? p->n : std::min(std::max<int64_t>(src_size, 0), p->n);
        if (is_pow2(dp) && is_pow2(sp) &&
            (dst_dist < footprint(p->n, dp) || src_dist < footprint(src_n, sp)))
            return fail(PCX_ERR_INVALID_ARGUMENT, "batch distance is smaller than one vector");
        if ((const T*)dst == src && dst_dist != src_dist)
            return fail(PCX_ERR_INVALID_ARGUMENT, "in-place batch needs equal source and destination distances");
    }
    return run_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist,
                     inverse ? p->n : src_size, batch, (hipStream_t)stream);
}
} // namespace

extern "C" int pcx_fft_batch(pcx_plan_t p, int inverse, int normalized, float* dst, int64_t dst_pack,
                             int64_t dst_dist, const float* src, int64_t src_pack, int64_t src_dist,
                             int64_t src_size, int64_t batch, void* stream) {
    return fft_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist, src_size, batch,
                     stream);
}

// ---------------------------------------------------------------- FP64 transforms
extern "C" int pcx_fft_f64(pcx_plan_t p, double* dst, int64_t dst_pack, const double* src, int64_t src_pack,
                           int64_t src_size, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 0, 0, dst, dst_pack, 0, src, src_pack, 0, src_size, 1, (hipStream_t)stream);
}

extern "C" int pcx_ifft_f64(pcx_plan_t p, int normalized, double* dst, int64_t dst_pack, const double* src,
                            int64_t src_pack, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 1, normalized, dst, dst_pack, 0, src, src_pack, 0, p->n, 1, (hipStream_t)stream);
}

extern "C" int pcx_fft_batch_f64(pcx_plan_t p, int inverse, int normalized, double* dst, int64_t dst_pack,
                                 int64_t dst_dist, const double* src, int64_t src_pack, int64_t src_dist,
                                 int64_t src_size, int64_t batch, void* stream) {
    return fft_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist, src_size, batch,
                     stream);
}

// ---------------------------------------------------------------- cross-correlation
struct pcx_fxcorr_s {
    int64_t n = 0, m = 0;
    int     lg = 0;
    int     precision = PCX_PREC_FP32;
    float2* tw = nullptr;
    float*  kspec = nullptr; // FFT(g)/N, split re[N] im[N]
    float*  halo  = nullptr;
    int64_t halo_cap = 0;
    std::mutex mu;
};

extern "C" int pcx_fxcorr_create(const float* g, int64_t g_pack, int64_t g_size, int64_t fft_size, void* stream,
                                 pcx_fxcorr_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "output pointer is null");
    *out = nullptr;
    if (!g)
        return fail(PCX_ERR_INVALID_ARGUMENT, "kernel pointer is null");
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_pack(g_pack, "kernel"))
        return e;
    if (g_size < 1 || g_size > fft_size)
        return fail(PCX_ERR_SIZE_MISMATCH,
                    "source size (which is " + s(g_size) + ") is bigger than fft size (which is " + s(fft_size) + ")");
    int lg = ilog2(fft_size);
    if (lg > kMaxSinglePass)
        return fail(PCX_ERR_UNSUPPORTED, "fxcorr fft_size (which is " + s(fft_size) + ") exceeds 16384");
    hipStream_t st = (hipStream_t)stream;
    auto*       x  = new pcx_fxcorr_s;
    x->n = fft_size, x->m = g_size, x->lg = lg;
    hipError_t he = hipMalloc(&x->tw, sizeof(float2) * fft_size);
    if (he == hipSuccess)
        he = hipMalloc(&x->kspec, sizeof(float) * 2 * fft_size);
    if (he == hipSuccess)
        he = gen_twiddles(x->tw, fft_size, 1, fft_size);
    if (he == hipSuccess)
        he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        pcx_fxcorr_destroy(x);
        return fail(PCX_ERR_HIP, std::string("fxcorr setup: ") + hipGetErrorString(he));
    }
    // kernel spectrum: FFT of g zero-extended to N, scaled by 1/N (fxcorr.hpp:67-68)
    RowArgs a{};
    a.src = g, a.dst = x->kspec, a.src_dist = 0, a.dst_dist = 0, a.src_size = g_size, a.nrows = 1;
    a.src_lp = ilog2(g_pack), a.dst_lp = lg, a.rows_lg = 0;
    a.order = 0, a.conj_in = 0, a.conj_out = 0, a.scale = (float)(1.0 / (double)fft_size);
    a.twA = x->tw;
    kSingle[lg - 1][0](a, 1, 0, st);
    he = hipGetLastError();
    if (he == hipSuccess)
        he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
        pcx_fxcorr_destroy(x);
        return fail(PCX_ERR_HIP, std::string("fxcorr kernel spectrum: ") + hipGetErrorString(he));
    }
    *out = x;
    return PCX_SUCCESS;
}

extern "C" int pcx_fxcorr_destroy(pcx_fxcorr_t x) {
    if (!x)
        return PCX_SUCCESS;
    (void)hipFree(x->tw);
    (void)hipFree(x->kspec);
    (void)hipFree(x->halo);
    delete x;
    return PCX_SUCCESS;
}

extern "C" int64_t pcx_fxcorr_step(pcx_fxcorr_t x, int64_t src_pack) {
    if (!x || !is_pow2(src_pack))
        return -1;
    int64_t overlap = x->m - 1;
    if (x->n <= overlap)
        return -1;
    // fxcorr.hpp:110: step = ((N - overlap) / P) * P
    return (x->n - overlap) / src_pack * src_pack;
}

extern "C" int pcx_fxcorr_set_precision(pcx_fxcorr_t x, int precision) {
    if (!x)
        return fail(PCX_ERR_INVALID_ARGUMENT, "fxcorr handle is null");
    if (precision != PCX_PREC_FP32 && precision != PCX_PREC_BF16X3 && precision != PCX_PREC_FP16X2)
        return fail(PCX_ERR_UNSUPPORTED, "precision " + s(precision) + " is not implemented");
    if (precision != PCX_PREC_FP32 && x->lg < 4)
        return fail(PCX_ERR_UNSUPPORTED, "split-precision correlation needs fft_size >= 16");
    x->precision = precision;
    return PCX_SUCCESS;
}

extern "C" int pcx_fxcorr_exec_batch(pcx_fxcorr_t x, float* dst, int64_t dst_pack, int64_t dst_dist, const float* src,
                                     int64_t src_pack, int64_t src_dist, int64_t len, int64_t channels, void* stream) {
    if (!x)
        return fail(PCX_ERR_INVALID_ARGUMENT, "fxcorr handle is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    if (len < 0 || channels < 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "signal length (which is " + s(len) + ") or channel count is negative");
    int64_t step = pcx_fxcorr_step(x, src_pack);
    if (step <= 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "fft size (which is " + s(x->n) + ") leaves no output per block for a kernel of " +
                                               s(x->m) + " taps at pack " + s(src_pack));
    const bool inplace = (const float*)dst == src;
    if (channels > 1 && (src_dist < footprint(len, src_pack) || dst_dist < footprint(len, dst_pack)))
        return fail(PCX_ERR_INVALID_ARGUMENT, "channel distance is smaller than one signal");
    if (inplace && (dst_pack != src_pack || dst_dist != src_dist))
        return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                s(src_pack) + ") to destination pack size (which is " + s(dst_pack) + ")");
    if (len == 0 || channels == 0)
        return PCX_SUCCESS;
    hipStream_t st      = (hipStream_t)stream;
    int64_t     nblocks = (len + step - 1) / step;
    XcArgs      a{};
    a.src = src, a.dst = dst, a.kre = x->kspec, a.kim = x->kspec + x->n, a.twA = x->tw;
    a.len = len, a.step = step, a.nblocks = nblocks, a.hlen = x->n - step;
    a.src_lp = ilog2(src_pack), a.dst_lp = ilog2(dst_pack);
    a.src_dist = src_dist, a.dst_dist = dst_dist, a.channels = channels;
    std::lock_guard<std::mutex> lock(x->mu);
    if (inplace && nblocks > 1 && a.hlen > 0) {
        // blocks overwrite the overlap that the next block still reads: save it first
        int64_t need = nblocks * channels * a.hlen;
        if (x->halo_cap < need) {
            if (x->halo) {
                HIPCHK(hipStreamSynchronize(st));
                HIPCHK(hipFree(x->halo));
                x->halo = nullptr;
            }
            hipError_t he = hipMalloc(&x->halo, sizeof(float) * 2 * need);
            if (he != hipSuccess)
                return fail(PCX_ERR_OUT_OF_MEMORY, "halo allocation failed");
            x->halo_cap = need;
        }
        a.hre = x->halo, a.him = x->halo + need;
        int64_t nb = std::min<int64_t>((need + 255) / 256, (int64_t)1 << 20);
        hipLaunchKernelGGL(k_fxcorr_halo, dim3((unsigned)nb), dim3(256), 0, st, src, a.src_lp, src_dist, len, step, a.hlen,
                           nblocks, channels, x->halo, x->halo + need);
        HIPCHK(hipGetLastError());
    }
    if (x->precision != PCX_PREC_FP32)
        kXcMfma[x->precision - 1][x->lg - 4](a, st);
    else
        kXc[x->lg - 1](a, st);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}

extern "C" int pcx_fxcorr_exec(pcx_fxcorr_t x, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
                               int64_t len, void* stream) {
    return pcx_fxcorr_exec_batch(x, dst, dst_pack, 0, src, src_pack, 0, len, 1, stream);
}

// ---------------------------------------------------------------- memory and utilities
extern "C" int pcx_malloc(void** ptr, size_t bytes) {
    if (!ptr)
        return fail(PCX_ERR_INVALID_ARGUMENT, "pointer is null");
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess)
        return fail(PCX_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    return PCX_SUCCESS;
}
extern "C" int pcx_free(void* ptr) {
    HIPCHK(hipFree(ptr));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return PCX_SUCCESS;
}
extern "C" int pcx_memset(void* dst, int value, size_t bytes) {
    HIPCHK(hipMemset(dst, value, bytes));
    return PCX_SUCCESS;
}
extern "C" int pcx_synchronize(void* stream) {
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return PCX_SUCCESS;
}
extern "C" int pcx_generate(float* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma, void* stream) {
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(pack, "destination"))
        return e;
    if (n <= 0)
        return PCX_SUCCESS;
    int64_t nb = std::min<int64_t>((n + 255) / 256, (int64_t)1 << 20);
    hipLaunchKernelGGL(k_generate<float>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, dst, ilog2(pack), n, seed,
                       dist, sigma);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
extern "C" int pcx_generate_f64(double* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma,
                                void* stream) {
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(pack, "destination"))
        return e;
    if (n <= 0)
        return PCX_SUCCESS;
    int64_t nb = std::min<int64_t>((n + 255) / 256, (int64_t)1 << 20);
    hipLaunchKernelGGL(k_generate<double>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, dst, ilog2(pack), n,
                       seed, dist, sigma);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}

// ---------------------------------------------------------------- uneven-batch work queue
namespace {

struct QGroup { // transforms of one size and direction outside the persistent kernel
    int      lg   = 0;
    int      inv  = 0;
    int64_t  count = 0;
    int64_t* src_offs = nullptr; // device
    int64_t* dst_offs = nullptr;
    pcx_plan_t plan = nullptr;
};

} // namespace

struct pcx_queue_s {
    int                 normalized = 1;
    int64_t             njobs      = 0;
    double              flops = 0, bytes = 0;
    // persistent part
    int32_t             nitems    = 0;
    QItem*              items     = nullptr; // device
    int64_t*            q_src_offs = nullptr;
    int64_t*            q_dst_offs = nullptr;
    uint32_t*           counter   = nullptr;
    float2*             tw        = nullptr; // W_A for A = 2 .. 4096
    int                 grid      = 1024;
    // grouped part
    std::vector<QGroup> groups;
    // captured graph
    hipGraph_t          graph = nullptr;
    hipGraphExec_t      exec  = nullptr;
};

namespace {

int queue_run(pcx_queue_t q, float* dst, const float* src, hipStream_t st) {
    if (q->nitems > 0) {
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, q->counter);
        QueueArgs a{};
        a.src = src, a.dst = dst, a.src_offs = q->q_src_offs, a.dst_offs = q->q_dst_offs;
        a.items = q->items, a.nitems = q->nitems, a.normalized = q->normalized, a.counter = q->counter, a.tw = q->tw;
        int grid = q->grid < q->nitems ? q->grid : q->nitems;
        hipLaunchKernelGGL(k_queue, dim3(grid), dim3(kQueueThreads), 0, st, a);
        HIPCHK(hipGetLastError());
    }
    for (auto& g : q->groups) {
        int64_t n = (int64_t)1 << g.lg;
        if (int e = run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st, g.src_offs,
                              g.dst_offs))
            return e;
    }
    return PCX_SUCCESS;
}

void queue_free(pcx_queue_t q) {
    if (q->exec)
        (void)hipGraphExecDestroy(q->exec);
    if (q->graph)
        (void)hipGraphDestroy(q->graph);
    (void)hipFree(q->items);
    (void)hipFree(q->q_src_offs);
    (void)hipFree(q->q_dst_offs);
    (void)hipFree(q->counter);
    (void)hipFree(q->tw);
    for (auto& g : q->groups) {
        (void)hipFree(g.src_offs);
        (void)hipFree(g.dst_offs);
        pcx_plan_destroy(g.plan);
    }
    delete q;
}

} // namespace

extern "C" int pcx_queue_create(const pcx_job* jobs, int64_t njobs, int normalized, pcx_queue_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue output pointer is null");
    *out = nullptr;
    if (njobs < 0 || (njobs > 0 && !jobs))
        return fail(PCX_ERR_INVALID_ARGUMENT, "job list is null or negative");
    for (int64_t i = 0; i < njobs; ++i) {
        if (jobs[i].log2n < 1 || jobs[i].log2n > kMaxLog2)
            return fail(PCX_ERR_INVALID_SIZE, "fft_size (which is  2^" + s(jobs[i].log2n) + ") is not supported");
        if (jobs[i].inverse != 0 && jobs[i].inverse != 1)
            return fail(PCX_ERR_INVALID_ARGUMENT, "job " + s(i) + ": inverse must be 0 or 1");
        if (jobs[i].src_offset < 0 || jobs[i].dst_offset < 0)
            return fail(PCX_ERR_INVALID_ARGUMENT, "job " + s(i) + ": negative offset");
    }
    auto* q       = new pcx_queue_s;
    q->normalized = normalized ? 1 : 0;
    q->njobs      = njobs;
    // bucket by (size, direction)
    std::vector<std::vector<int64_t>> bucket((kMaxLog2 + 1) * 2);
    for (int64_t i = 0; i < njobs; ++i) {
        bucket[jobs[i].log2n * 2 + jobs[i].inverse].push_back(i);
        double n = (double)((int64_t)1 << jobs[i].log2n);
        q->flops += 5.0 * n * (double)jobs[i].log2n;
        q->bytes += 16.0 * n;
    }
    std::vector<QItem>   items;
    std::vector<int64_t> qs, qd;
    hipError_t           he = hipSuccess;
    // persistent items, largest size first (longest-processing-time order)
    for (int lg = kQueueMaxLog2; lg >= 1; --lg) {
        int per = kQueueThreads / (lg >= 4 ? (1 << lg) / 16 : 1);
        for (int inv = 0; inv < 2; ++inv) {
            auto& b = bucket[lg * 2 + inv];
            for (size_t k = 0; k < b.size(); k += per) {
                int cnt = (int)std::min<size_t>(per, b.size() - k);
                items.push_back({lg, inv, (int32_t)qs.size(), cnt});
                for (int c = 0; c < cnt; ++c) {
                    qs.push_back(jobs[b[k + c]].src_offset);
                    qd.push_back(jobs[b[k + c]].dst_offset);
                }
            }
        }
    }
    q->nitems = (int32_t)items.size();
    if (!items.empty()) {
        he = hipMalloc(&q->items, sizeof(QItem) * items.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->q_src_offs, sizeof(int64_t) * qs.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->q_dst_offs, sizeof(int64_t) * qd.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->counter, sizeof(uint32_t));
        if (he == hipSuccess)
            he = hipMalloc(&q->tw, sizeof(float2) * ((1 << (kQueueMaxLog2 + 1)) - 2));
        if (he == hipSuccess)
            he = hipMemcpy(q->items, items.data(), sizeof(QItem) * items.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(q->q_src_offs, qs.data(), sizeof(int64_t) * qs.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(q->q_dst_offs, qd.data(), sizeof(int64_t) * qd.size(), hipMemcpyHostToDevice);
        for (int lg = 1; lg <= kQueueMaxLog2 && he == hipSuccess; ++lg)
            he = gen_twiddles(q->tw + ((1 << lg) - 2), (int64_t)1 << lg, 1, (int64_t)1 << lg);
    }
    // grouped launches for the larger sizes, largest first
    for (int lg = kMaxLog2; lg > kQueueMaxLog2 && he == hipSuccess; --lg) {
        for (int inv = 0; inv < 2; ++inv) {
            auto& b = bucket[lg * 2 + inv];
            if (b.empty())
                continue;
            QGroup g;
            g.lg = lg, g.inv = inv, g.count = (int64_t)b.size();
            std::vector<int64_t> so(b.size()), dof(b.size());
            for (size_t k = 0; k < b.size(); ++k)
                so[k] = jobs[b[k]].src_offset, dof[k] = jobs[b[k]].dst_offset;
            he = hipMalloc(&g.src_offs, sizeof(int64_t) * b.size());
            if (he == hipSuccess)
                he = hipMalloc(&g.dst_offs, sizeof(int64_t) * b.size());
            if (he == hipSuccess)
                he = hipMemcpy(g.src_offs, so.data(), sizeof(int64_t) * b.size(), hipMemcpyHostToDevice);
            if (he == hipSuccess)
                he = hipMemcpy(g.dst_offs, dof.data(), sizeof(int64_t) * b.size(), hipMemcpyHostToDevice);
            if (he == hipSuccess) {
                if (int e = pcx_plan_create((int64_t)1 << lg, 2048, PCX_ORDER_NORMAL, &g.plan)) {
                    q->groups.push_back(g);
                    queue_free(q);
                    return e;
                }
            }
            q->groups.push_back(g);
        }
    }
    if (he == hipSuccess)
        he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        queue_free(q);
        return fail(PCX_ERR_HIP, std::string("queue setup: ") + hipGetErrorString(he));
    }
    *out = q;
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_destroy(pcx_queue_t q) {
    if (q)
        queue_free(q);
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_exec(pcx_queue_t q, float* dst, const float* src, void* stream) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    return queue_run(q, dst, src, (hipStream_t)stream);
}

extern "C" int pcx_queue_capture(pcx_queue_t q, float* dst, const float* src, void* stream) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    hipStream_t st = (hipStream_t)stream;
    if (!st)
        return fail(PCX_ERR_INVALID_ARGUMENT, "graph capture needs a non-default stream");
    // make sure every group's scratch exists before capture (allocation is not capturable)
    for (auto& g : q->groups) {
        int64_t n = (int64_t)1 << g.lg;
        if (g.plan->two_pass && g.plan->scratch_cap < std::min(g.count, g.plan->chunk)) {
            if (int e = run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st,
                                  g.src_offs, g.dst_offs))
                return e;
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    if (q->exec) {
        HIPCHK(hipGraphExecDestroy(q->exec));
        q->exec = nullptr;
    }
    if (q->graph) {
        HIPCHK(hipGraphDestroy(q->graph));
        q->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int        e  = queue_run(q, dst, src, st);
    hipGraph_t gr = nullptr;
    hipError_t he = hipStreamEndCapture(st, &gr);
    if (e)
        return e;
    if (he != hipSuccess)
        return fail(PCX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    q->graph = gr;
    HIPCHK(hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0));
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_launch(pcx_queue_t q, void* stream) {
    if (!q || !q->exec)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue has no captured graph");
    HIPCHK(hipGraphLaunch(q->exec, (hipStream_t)stream));
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_work(pcx_queue_t q, double* flops, double* bytes) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (flops)
        *flops = q->flops;
    if (bytes)
        *bytes = q->bytes;
    return PCX_SUCCESS;
}

namespace {

// Times fn (which enqueues work on st) with events: best of `reps` after one warm-up.
template<class F>
float time_ms(hipStream_t st, F&& fn, int reps = 3) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess)
        return 1e30f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return 1e30f;
    }
    float best = 1e30f;
    if (fn() != PCX_SUCCESS)
        best = -1.f;
    for (int r = 0; r < reps && best >= 0.f; ++r) {
        float ms = 0.f;
        (void)hipEventRecord(a, st);
        if (fn() != PCX_SUCCESS) {
            best = -1.f;
            break;
        }
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        best = ms < best ? ms : best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best < 0.f ? 1e30f : best;
}

} // namespace

// Autotuner: persistent grid size, then per group the four-step shape and grid caps.  Winners are
// written to cache_path (one "key value..." line each) and read back from it when present.
extern "C" int pcx_queue_autotune(pcx_queue_t q, float* dst, const float* src, void* stream, const char* cache_path) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    hipStream_t st = (hipStream_t)stream;
    if (cache_path) { // reuse earlier decisions
        if (FILE* f = std::fopen(cache_path, "r")) {
            char key[32];
            int  a = 0, b = 0, c = 0, d = 0, e = 0;
            long long gc = 0, gr = 0;
            while (std::fscanf(f, "%31s", key) == 1) {
                if (!std::strcmp(key, "queue_grid") && std::fscanf(f, "%d", &a) == 1)
                    q->grid = a;
                else if (!std::strcmp(key, "group") && std::fscanf(f, "%d %d %d %d %d %lld %lld", &a, &b, &c, &d, &e, &gc, &gr) == 7) {
                    for (auto& g : q->groups)
                        if (g.lg == a && g.inv == b && g.plan->two_pass) {
                            if (const FourStep* sh = find_shape(a, c, d, e)) {
                                if (int err = plan_apply_shape(g.plan, sh)) {
                                    std::fclose(f);
                                    return err;
                                }
                                g.plan->grid_col = gc, g.plan->grid_row = gr;
                            }
                        }
                } else {
                    break;
                }
            }
            std::fclose(f);
            return PCX_SUCCESS;
        }
    }
    if (q->nitems > 0) {
        float best = 1e30f;
        int   bg   = q->grid;
        for (int grid : {256, 512, 1024, 2048, 4096}) {
            q->grid  = grid;
            float ms = time_ms(st, [&] {
                std::vector<QGroup> keep;
                keep.swap(q->groups); // time the persistent part alone
                int e = queue_run(q, dst, src, st);
                keep.swap(q->groups);
                return e;
            });
            if (ms < best)
                best = ms, bg = grid;
        }
        q->grid = bg;
    }
    for (auto& g : q->groups) {
        if (!g.plan->two_pass)
            continue;
        int64_t         n    = (int64_t)1 << g.lg;
        const FourStep* best = g.plan->shape;
        int64_t         bgc = g.plan->grid_col, bgr = g.plan->grid_row;
        float           bms = 1e30f;
        for (const auto& sh : kShapes) {
            if (sh.lg != g.lg)
                continue;
            if (int err = plan_apply_shape(g.plan, &sh))
                return err;
            for (int64_t gc : {(int64_t)0, (int64_t)1024, (int64_t)2048}) {
                g.plan->grid_col = g.plan->grid_row = gc;
                float ms = time_ms(st, [&] {
                    return run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st,
                                     g.src_offs, g.dst_offs);
                });
                if (ms < bms)
                    bms = ms, best = &sh, bgc = gc, bgr = gc;
            }
        }
        if (int err = plan_apply_shape(g.plan, best))
            return err;
        g.plan->grid_col = bgc, g.plan->grid_row = bgr;
    }
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "w")) {
            std::fprintf(f, "queue_grid %d\n", q->grid);
            for (auto& g : q->groups)
                if (g.plan->two_pass)
                    std::fprintf(f, "group %d %d %d %d %d %lld %lld\n", g.lg, g.inv, g.plan->shape->n2, g.plan->shape->g,
                                 g.plan->shape->c, (long long)g.plan->grid_col, (long long)g.plan->grid_row);
            std::fclose(f);
        }
    }
    return PCX_SUCCESS;
}

// ---------------------------------------------------------------- vector expressions
namespace {
// One validation and launch path for both element types (S, I: the C ABI's source and
// instruction structs of that type).
template<class T, class S, class I>
int vexpr_eval(const S* srcs, int32_t nsrc, const I* prog, int32_t nprog, T* dst, int64_t dst_pack, int64_t n,
               void* stream) {
    if (nsrc < 0 || nsrc > PCX_VEXPR_MAX_SRC || (nsrc > 0 && !srcs))
        return fail(PCX_ERR_INVALID_ARGUMENT, "source count (which is " + s(nsrc) + ") is not in [0, 16]");
    if (nprog < 1 || nprog > PCX_VEXPR_MAX_PROG || !prog)
        return fail(PCX_ERR_INVALID_ARGUMENT, "program length (which is " + s(nprog) + ") is not in [1, 64]");
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "destination pointer is null");
    if (n < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "size (which is " + s(n) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    VExprArgsT<T> a{};
    for (int k = 0; k < nsrc; ++k) {
        if (!srcs[k].data)
            return fail(PCX_ERR_INVALID_ARGUMENT, "source " + s(k) + " is null");
        if (int e = check_pack(srcs[k].pack, "source"))
            return e;
        if ((const T*)dst == srcs[k].data && srcs[k].pack != dst_pack)
            return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                    s(srcs[k].pack) + ") to destination pack size (which is " +
                                                    s(dst_pack) + ")");
        a.src[k].data = srcs[k].data;
        a.src[k].lp   = ilog2(srcs[k].pack);
    }
    for (int k = 0; k < nprog; ++k) {
        const I& in = prog[k];
        bool     ok = in.op >= 0 && in.op < VOP_COUNT && in.d >= 0 && in.d < PCX_VEXPR_REGS;
        if (in.op == VOP_LOAD)
            ok = ok && in.a >= 0 && in.a < nsrc;
        else
            ok = ok && in.a >= 0 && in.a < PCX_VEXPR_REGS && (in.b >= 0 && in.b < PCX_VEXPR_REGS);
        if (!ok)
            return fail(PCX_ERR_INVALID_ARGUMENT, "instruction " + s(k) + " is malformed");
        a.prog[k] = VInstrT<T>{in.op, in.d, in.a, in.b, in.s_re, in.s_im};
    }
    a.nprog  = nprog;
    a.dst    = dst;
    a.dst_lp = ilog2(dst_pack);
    a.n      = n;
    if (n == 0)
        return PCX_SUCCESS;
    int64_t nwg = (n + 255) / 256;
    if (nwg > 16384)
        nwg = 16384;
    hipLaunchKernelGGL(k_vexpr<T>, dim3((unsigned)nwg), dim3(256), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
} // namespace

extern "C" int pcx_vexpr_eval(const pcx_vsource* srcs, int32_t nsrc, const pcx_vinstr* prog, int32_t nprog, float* dst,
                              int64_t dst_pack, int64_t n, void* stream) {
    return vexpr_eval(srcs, nsrc, prog, nprog, dst, dst_pack, n, stream);
}

extern "C" int pcx_repack(float* dst, int64_t dst_pack, const float* src, int64_t src_pack, int64_t n, void* stream) {
    pcx_vsource sv{src, src_pack};
    pcx_vinstr  ld{PCX_VOP_LOAD, 0, 0, 0, 0.f, 0.f};
    return pcx_vexpr_eval(&sv, 1, &ld, 1, dst, dst_pack, n, stream);
}

extern "C" int pcx_vexpr_eval_f64(const pcx_vsource_f64* srcs, int32_t nsrc, const pcx_vinstr_f64* prog, int32_t nprog,
                                  double* dst, int64_t dst_pack, int64_t n, void* stream) {
    return vexpr_eval(srcs, nsrc, prog, nprog, dst, dst_pack, n, stream);
}

extern "C" int pcx_repack_f64(double* dst, int64_t dst_pack, const double* src, int64_t src_pack, int64_t n,
                              void* stream) {
    pcx_vsource_f64 sv{src, src_pack};
    pcx_vinstr_f64  ld{PCX_VOP_LOAD, 0, 0, 0, 0.0, 0.0};
    return pcx_vexpr_eval_f64(&sv, 1, &ld, 1, dst, dst_pack, n, stream);
}

// ---------------------------------------------------------------- FFT across vectors
namespace {
using AcrossLaunch = void (*)(const AcrossArgs&, int64_t, hipStream_t);
template<int LA>
void launch_across(const AcrossArgs& a, int64_t grid, hipStream_t st) {
    constexpr int C = LA <= 8 ? 64 : (LA == 9 ? 32 : 16);
    int64_t       n = (a.len + C - 1) / C;
    if (n > grid)
        n = grid;
    hipLaunchKernelGGL((k_across<LA, C>), dim3((unsigned)fit_grid(grid8(n), C * Plan<LA>::TA)), dim3(C * Plan<LA>::TA), 0,
                       st, a);
}
template<int... L>
constexpr auto across_table(std::integer_sequence<int, L...>) {
    return std::array<AcrossLaunch, sizeof...(L)>{&launch_across<L + 1>...};
}
constexpr int kMaxAcross = 10;
const auto    kAcross    = across_table(std::make_integer_sequence<int, kMaxAcross>{});
} // namespace

extern "C" int pcx_fft_across(pcx_plan_t p, int inverse, int normalized, float* const* dst_vecs, int64_t dst_pack,
                              const float* const* src_vecs, int64_t src_pack, int64_t length, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64)
        return fail(PCX_ERR_UNSUPPORTED, "transforms across vectors are built for float plans");
    if (!dst_vecs || !src_vecs)
        return fail(PCX_ERR_INVALID_ARGUMENT, "vector table is null");
    if (length < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "length (which is " + s(length) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    if (p->order != PCX_ORDER_NORMAL)
        return fail(PCX_ERR_UNSUPPORTED, "transforms across vectors are built for the normal order");
    if (p->lg > kMaxAcross)
        return fail(PCX_ERR_UNSUPPORTED, "transforms across vectors cover at most 1024 vectors (fft_size " + s(p->n) + ")");
    if (length == 0)
        return PCX_SUCCESS;
    AcrossArgs a{};
    a.src = src_vecs, a.dst = dst_vecs, a.len = length;
    a.src_lp = ilog2(src_pack), a.dst_lp = ilog2(dst_pack);
    a.conj_in = inverse, a.conj_out = inverse;
    a.scale   = (inverse && normalized) ? (float)(1.0 / (double)p->n) : 1.0f;
    a.twA     = p->tw_row;
    kAcross[p->lg - 1](a, 2048, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
