// k_f64.hip -- FP64 kernels: single pass (N <= 8192, normal and bit_reversed order) and the
// default four-step shapes 2^14 .. 2^23.
#include "shapes.hpp"

namespace pcxd {
namespace {

// ---- fft_unit<double, Order>: natural and bit_reversed orders (the reference has no FP64
// unordered kernel: avx2_fft.hpp:114-217 is written for float only)
template<int LA>
constexpr std::array<RowLaunchT<double>, 5> single_set64() {
    using D            = double;
    constexpr int  G   = default_rows<LA, D>();
    constexpr bool UNI = Plan<LA, false, D>::TA >= 64;
    return {&launch_row<LA, G, LD_CONTIG, ST_CONTIG, 0, UNI, 0, PREC_FP32, D>,
            &launch_row<LA, G, LD_CONTIG, ST_PERM, 0, UNI, ORD_BITREV, PREC_FP32, D>, nullptr,
            &launch_row<LA, G, LD_PERM, ST_CONTIG, 0, UNI, ORD_BITREV, PREC_FP32, D>, nullptr};
}
template<int... L>
constexpr auto single_table64(std::integer_sequence<int, L...>) {
    return std::array<std::array<RowLaunchT<double>, 5>, sizeof...(L)>{single_set64<L + 1>()...};
}

// Defaults (first match per size); forward ms for a 2 GiB batch measured on MI355X.
// 2^14 .. 2^16: one workgroup per CU with 512-byte column segments (2^14, 2^15: 64 columns) and
// 256-byte segments in both passes at 2^16 (32 columns, 32 rows): 1.63 / 1.58 / 1.62 ms against
// 1.68 / 1.68 / 1.73 ms with the two-workgroups-per-CU tiles of the alternatives.  Wider segments
// win over occupancy, e.g. 2^20 with C = G = 8 and 129-133 KiB per workgroup runs in 2.43 ms
// against 3.23 ms for C = 8, G = 2.
constexpr FourStepT<double> kDefault64[] = {
    shape64<14, 7, 7, 32, 64>(), shape64<15, 8, 7, 32, 64>(), shape64<16, 8, 8, 32, 32>(),
    shape64<17, 8, 9, 32, 16>(),  // 1.84 ms
    shape64<18, 10, 8, 8, 32>(),  // 1.89
    shape64<19, 10, 9, 8, 16>(),  // 2.41
    shape64<20, 10, 10, 8, 8>(),  // 2.43
    shape64<21, 11, 10, 4, 8>(),  // 3.11
    shape64<22, 12, 10, 4, 8>(),  // 3.12
    shape64<23, 13, 10, 2, 4>(),  // 4.91
};

const std::array<std::array<RowLaunchT<double>, 5>, kMaxSinglePass64> kSingle64 =
    single_table64(std::make_integer_sequence<int, kMaxSinglePass64>{});

// Three-pass natural-order transforms of 2^21 .. 2^23 doubles: tiles of 8192 points (128 KiB),
// first-pass columns of 2^7 / 2^8 (64 / 32 per tile, 512- / 256-byte segments), 64 second-pass
// columns per tile and rows of 2^7 / 2^8 (64 / 32 per workgroup).
template<int LG, int N2, int NA>
constexpr ThreePassT<double> three_pass64() {
    using D          = double;
    constexpr int N1 = LG - N2, C1 = 8192 >> N2, G3 = 8192 >> NA;
    static_assert(N1 == NA + kThreeNbLog2, "three-pass split");
    return {LG, N2, NA, C1, 64, &launch_col<N2, C1, tw_cap64<LG>(), PREC_FP32, D>,
            &launch_col<kThreeNbLog2, 64, tw_cap64<N1>(), PREC_FP32, D>,
            &launch_row<NA, G3, LD_CONTIG, ST_STRIDED, 0, true, 0, PREC_FP32, D>};
}
constexpr ThreePassT<double> kThree64[] = {three_pass64<21, 7, 7>(), three_pass64<22, 8, 7>(),
                                           three_pass64<23, 8, 8>()};

} // namespace

const std::array<std::array<RowLaunchT<double>, 5>, kMaxSinglePass64>& single64_tab() { return kSingle64; }
ShapeList<double> shapes64_default() { return {kDefault64, (int)(sizeof(kDefault64) / sizeof(kDefault64[0]))}; }
ThreeList<double> three_pass64_tab() { return {kThree64, (int)(sizeof(kThree64) / sizeof(kThree64[0]))}; }

} // namespace pcxd
