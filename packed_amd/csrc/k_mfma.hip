// k_mfma.hip -- split-precision (BF16X3, FP16X2) matrix-core variants: single pass 2^4 .. 2^14,
// correlation blocks 2^4 .. 2^14 and the default four-step shapes of 2^15 .. 2^20.
#include "dispatch.hpp"

namespace pcxd {
namespace {

template<int PREC, int... L>
constexpr auto single_mfma_table(std::integer_sequence<int, L...>) {
    return std::array<RowLaunch, sizeof...(L)>{
        &launch_row<L + 4, default_rows<L + 4>(), LD_CONTIG, ST_CONTIG, 0, (Plan<L + 4>::TA >= 64), 0, PREC>...};
}
template<int PREC, int... L>
constexpr auto xc_mfma_table(std::integer_sequence<int, L...>) {
    return std::array<XcLaunch, sizeof...(L)>{&launch_xc<L + 4, PREC>...};
}
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


const std::array<std::array<RowLaunch, kMaxSinglePass - 3>, 2> kSingleMfma = {
    single_mfma_table<PREC_BF16X3>(std::make_integer_sequence<int, kMaxSinglePass - 3>{}),
    single_mfma_table<PREC_FP16X2>(std::make_integer_sequence<int, kMaxSinglePass - 3>{})};
const std::array<std::array<XcLaunch, kMaxSinglePass - 3>, 2> kXcMfma = {
    xc_mfma_table<PREC_BF16X3>(std::make_integer_sequence<int, kMaxSinglePass - 3>{}),
    xc_mfma_table<PREC_FP16X2>(std::make_integer_sequence<int, kMaxSinglePass - 3>{})};
// 128-column tiles at 2^15 and 2^16 as in the FP32 defaults (2^16 x 4096 BF16X3 / FP16X2:
// 1.93 / 1.84 ms per step against 2.00 / 1.92 with 64 columns)
const std::array<FourStepMfma, 6> kShapesMfma = {shape_mfma<15, 8, 7, 32, 128>(), shape_mfma<16, 9, 7, 32, 128>(),
                                                 shape_mfma<17, 9, 8, 32, 64>(),  shape_mfma<18, 9, 9, 32, 32>(),
                                                 shape_mfma<19, 9, 10, 32, 16>(), shape_mfma<20, 10, 10, 16, 16>()};

} // namespace

const std::array<std::array<RowLaunch, kMaxSinglePass - 3>, 2>& single_mfma_tab() { return kSingleMfma; }
const std::array<std::array<XcLaunch, kMaxSinglePass - 3>, 2>&  xc_mfma_tab() { return kXcMfma; }
const std::array<FourStepMfma, 6>&                              shapes_mfma() { return kShapesMfma; }

} // namespace pcxd
