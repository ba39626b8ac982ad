// k_single.hip -- single-pass FP32 kernels (N <= 16384, every order), the one-pass 2^15
// decimation-in-frequency kernel and the transform across vectors.
#include "dispatch.hpp"

namespace pcxd {
namespace {

template<int LA>
constexpr RowSet single_set() {
    constexpr int  G   = default_rows<LA>();
    constexpr bool UNI = Plan<LA>::TA >= 64;
    return {&launch_row<LA, G, LD_CONTIG, ST_CONTIG, 0, UNI, 0>, &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_BITREV>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_UNORD>, &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_BITREV>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_UNORD>};
}
template<int... L>
constexpr auto single_table(std::integer_sequence<int, L...>) {
    return std::array<RowSet, sizeof...(L)>{single_set<L + 1>()...};
}

// One-pass decimation-in-frequency kernel (k_dif) for N = 2^15 = M * R: natural order with
// M = 2^14, R = 2; bit_reversed / unordered with M = 2^13, R = 4 (two workgroups per CU; the
// stores are contiguous there).  Measured on MI355X against the four-step for a 2 GiB batch:
// 1.55 vs 1.72 ms (natural), 1.50 vs 1.70 ms (bit_reversed).  For 2^16 and 2^17 the R-fold
// rereads cost more than the second pass saves (2.27 vs 1.71 ms at 2^16), so those sizes stay
// four-step.
template<int LM, int LR, int ORD>
void launch_dif(const RowArgs& a, int64_t batch, hipStream_t st) {
    int64_t nwg = ((batch + 7) / 8) * 8 * ((int64_t)1 << LR);
    hipLaunchKernelGGL((k_dif<LM, LR, ORD>), dim3((unsigned)nwg), dim3(Plan<LM, false, float, dif_ppt<LM>()>::TA), 0, st,
                       a);
}

// Columns per tile of the transform across 2^LA vectors: 128 KiB of LDS at most (C consecutive
// elements of each vector per wavefront read).
template<int LA, class T>
constexpr int across_cols() {
    constexpr int c = (65536 / (int)sizeof(T)) >> LA; // 2 * 2^LA * C * sizeof(T) <= 128 KiB
    constexpr int m = 64 * 4 / (int)sizeof(T);
    return c > m ? m : (c < 1 ? 1 : c);
}
template<int LA, class T>
void launch_across(const AcrossArgsT<T>& a, int64_t grid, hipStream_t st) {
    constexpr int C  = across_cols<LA, T>();
    constexpr int NT = C * Plan<LA, false, T>::TA;
    static_assert(NT <= 1024, "across tile");
    int64_t n = (a.len + C - 1) / C;
    if (n > grid)
        n = grid;
    hipLaunchKernelGGL((k_across<LA, C, T>), dim3((unsigned)fit_grid(grid8(n), NT)), dim3(NT), 0, st, a);
}
template<class T, int... L>
constexpr auto across_table(std::integer_sequence<int, L...>) {
    return std::array<void (*)(const AcrossArgsT<T>&, int64_t, hipStream_t), sizeof...(L)>{
        &launch_across<L + 1, T>...};
}

const std::array<RowSet, kMaxSinglePass> kSingle = single_table(std::make_integer_sequence<int, kMaxSinglePass>{});
const std::array<DifLaunch, 3> kDif = {&launch_dif<14, 1, ORD_NORMAL>, &launch_dif<13, 2, ORD_BITREV>,
                                       &launch_dif<13, 2, ORD_UNORD>};
const std::array<AcrossLaunch, kMaxAcross> kAcross =
    across_table<float>(std::make_integer_sequence<int, kMaxAcross>{});
const std::array<AcrossLaunch64, kMaxAcross64> kAcross64 =
    across_table<double>(std::make_integer_sequence<int, kMaxAcross64>{});

} // namespace

const std::array<RowSet, kMaxSinglePass>& single_tab() { return kSingle; }
const std::array<DifLaunch, 3>&            dif_tab() { return kDif; }
const std::array<AcrossLaunch, kMaxAcross>& across_tab() { return kAcross; }
const std::array<AcrossLaunch64, kMaxAcross64>& across64_tab() { return kAcross64; }

} // namespace pcxd
