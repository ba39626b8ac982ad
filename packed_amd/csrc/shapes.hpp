// shapes.hpp -- builders of the four-step shape entries (included by the k_shapes*.hip units).
#pragma once
#include "dispatch.hpp"

namespace pcxd {

// The G rows of a row workgroup lie in one transform (G <= N2), so buffer addressing applies.
template<int LA, int G, int LG>
constexpr std::array<RowLaunch, 6> pass_set() {
    constexpr int TW = tw_cap<LG>();
    return {&launch_row<LA, G, LD_CONTIG, ST_STRIDED, 0, true, 0>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, true, ORD_BITREV>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, true, ORD_UNORD>,
            &launch_row<LA, G, LD_STRIDED, ST_CONTIG, TW, true, 0>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, TW, true, ORD_BITREV>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, TW, true, ORD_UNORD>};
}
// PX: plan options of the columns (points per thread, one-array exchange).
template<int LA, int C, int LG, int PX = 0>
constexpr std::array<ColLaunch, 2> col_set() {
    return {&launch_col<LA, C, 0, PREC_FP32, float, PX>, &launch_col<LA, C, tw_cap<LG>(), PREC_FP32, float, PX>};
}
template<int LG, int N1, int N2, int G, int C, int PXC = 0>
constexpr FourStep shape() {
    static_assert(N1 + N2 == LG, "shape");
    return {LG, N1, N2, G, C, pass_set<N1, G, LG>(), col_set<N2, C, LG, PXC>()};
}

// FP64 shapes, 256-512 threads per workgroup.  C and G are the doubles per column segment and
// per row store segment.  Pass-twiddle tables above 512 entries are read from global memory.
template<int LG, int N1, int N2, int G, int C>
constexpr FourStepT<double> shape64() {
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

} // namespace pcxd
