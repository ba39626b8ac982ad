// k_mfma_big.hip -- split-precision (BF16X3, FP16X2) four-step kernels for 2^21 .. 2^24: the
// default FP32 shapes' splits and tiles (k_shapes.hip) with matrix-core butterflies, every order
// and both directions.  (The FP32 natural order of these sizes runs in three passes; the split
// forms keep the four-step, whose row kernels carry the permuted orders.)
#include "shapes.hpp"

namespace pcxd {
namespace {

template<int LG, int N1, int N2, int G, int C, int PREC>
constexpr FourStep shape_split() {
    static_assert(N1 + N2 == LG, "shape");
    constexpr int TW = tw_cap<LG>();
    return {LG,
            N1,
            N2,
            G,
            C,
            {&launch_row<N1, G, LD_CONTIG, ST_STRIDED, 0, true, 0, PREC>,
             &launch_row<N1, G, LD_CONTIG, ST_PERM, 0, true, ORD_BITREV, PREC>,
             &launch_row<N1, G, LD_CONTIG, ST_PERM, 0, true, ORD_UNORD, PREC>,
             &launch_row<N1, G, LD_STRIDED, ST_CONTIG, TW, true, 0, PREC>,
             &launch_row<N1, G, LD_PERM, ST_CONTIG, TW, true, ORD_BITREV, PREC>,
             &launch_row<N1, G, LD_PERM, ST_CONTIG, TW, true, ORD_UNORD, PREC>},
            {&launch_col<N2, C, 0, PREC>, &launch_col<N2, C, TW, PREC>}};
}
template<int PREC>
constexpr std::array<FourStep, 4> big_set() {
    return {shape_split<21, 11, 10, 8, 16, PREC>(), shape_split<22, 12, 10, 4, 16, PREC>(),
            shape_split<23, 12, 11, 4, 8, PREC>(), shape_split<24, 14, 10, 2, 16, PREC>()};
}

const MfmaBig kBig = {big_set<PREC_BF16X3>(), big_set<PREC_FP16X2>()};

} // namespace

const MfmaBig& shapes_mfma_big() { return kBig; }

} // namespace pcxd
