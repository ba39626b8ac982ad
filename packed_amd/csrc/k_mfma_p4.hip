// k_mfma_p4.hip -- split-precision (BF16X3, FP16X2) row kernels of the four-step shapes 2^15 ..
// 2^20 for the bit_reversed and unordered orders: the forward row pass stores permuted, the
// inverse row pass loads permuted (the column kernels do not depend on the order).  The rows
// and columns of the natural order are in k_mfma.hip, whose shape list this table follows.
#include "dispatch.hpp"

namespace pcxd {
namespace {

template<int LG, int N1, int G, int PREC>
constexpr std::array<RowLaunch, 4> perm_rows() {
    constexpr int TW = tw_cap<LG>();
    return {&launch_row<N1, G, LD_CONTIG, ST_PERM, 0, true, ORD_BITREV, PREC>,
            &launch_row<N1, G, LD_CONTIG, ST_PERM, 0, true, ORD_UNORD, PREC>,
            &launch_row<N1, G, LD_PERM, ST_CONTIG, TW, true, ORD_BITREV, PREC>,
            &launch_row<N1, G, LD_PERM, ST_CONTIG, TW, true, ORD_UNORD, PREC>};
}
template<int LG, int N1, int N2, int G>
constexpr std::array<std::array<RowLaunch, 4>, 2> perm_shape() {
    static_assert(N1 + N2 == LG, "shape");
    return {perm_rows<LG, N1, G, PREC_BF16X3>(), perm_rows<LG, N1, G, PREC_FP16X2>()};
}

// the same order and splits as shapes_mfma()
const MfmaPermRows kPermRows = {perm_shape<15, 8, 7, 32>(), perm_shape<16, 9, 7, 32>(), perm_shape<17, 9, 8, 32>(),
                                perm_shape<18, 9, 9, 32>(), perm_shape<19, 9, 10, 32>(), perm_shape<20, 10, 10, 16>()};

} // namespace

const MfmaPermRows& shapes_mfma_perm() { return kPermRows; }

} // namespace pcxd
