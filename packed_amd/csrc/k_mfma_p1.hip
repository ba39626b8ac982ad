// k_mfma_p1.hip -- split-precision (BF16X3, FP16X2) single-pass variants for the bit_reversed and
// unordered orders, 2^4 .. 2^14: permuted stores (forward) and permuted loads (inverse) around the
// matrix-core Stockham core.  The natural-order variants are in k_mfma.hip.
#include "dispatch.hpp"

namespace pcxd {
namespace {

template<int LA, int PREC>
constexpr RowSet mfma_perm_set() {
    constexpr int  G   = default_rows<LA>();
    constexpr bool UNI = Plan<LA>::TA >= 64;
    return {nullptr, &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_BITREV, PREC>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_UNORD, PREC>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_BITREV, PREC>,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_UNORD, PREC>};
}
template<int PREC, int... L>
constexpr auto mfma_perm_table(std::integer_sequence<int, L...>) {
    return std::array<RowSet, sizeof...(L)>{mfma_perm_set<L + 4, PREC>()...};
}

const MfmaSets kPerm = {mfma_perm_table<PREC_BF16X3>(std::make_integer_sequence<int, kMaxSinglePass - 3>{}),
                        mfma_perm_table<PREC_FP16X2>(std::make_integer_sequence<int, kMaxSinglePass - 3>{})};

} // namespace

const MfmaSets& single_mfma_perm_tab() { return kPerm; }

} // namespace pcxd
