// k_shapes.hip -- default FP32 four-step shapes, 2^15 .. 2^24 (first match per size).
#include "shapes.hpp"

namespace pcxd {
namespace {
constexpr FourStep kDefault[] = {
    // measured best on MI355X for 2^16 is N1 = 512, N2 = 128 with 128-column tiles (512-byte load
    // and store segments in the column pass; 128 KiB of LDS, one workgroup of 1024 threads per
    // CU): 0.778 ms per 2 GiB against 0.855 ms with 64 columns; 2^15 (four-step forward, 64 rows
    // per row workgroup) 1.59 against 1.70 ms per 2 GiB
    shape<15, 8, 7, 64, 128>(), shape<16, 9, 7, 32, 128>(), shape<17, 9, 8, 32, 64>(), shape<18, 9, 9, 32, 32>(),
    // 2^19: 512-point rows, 32 per workgroup (128-byte store segments), 1.91 against 2.03 ms per
    // 2 GiB with 1024-point rows (an alternative); 1024-point columns 32 per tile (128-byte
    // segments) with 32 points per thread and one LDS array: 1.83 / 1.83 ms forward / inverse
    // against 1.85 / 1.89 with 16 per tile.  (2^20 gained nothing from the same columns.)
    shape<19, 9, 10, 32, 32, 32 | kHalfExchange>(), shape<20, 10, 10, 16, 16>(), shape<21, 11, 10, 8, 16>(), shape<22, 12, 10, 4, 16>(),
    // 2^23, 2^24: more rows per workgroup than the LDS of two workgroups per CU allows (they
    // spill, but store 4 / 2 floats per segment instead of 2 / 1): 5.03 / 5.87 ms per 2 GiB
    // forward against 7.47 / 8.94
    shape<23, 12, 11, 4, 8>(), shape<24, 14, 10, 2, 16>(),
};
// N2 and Na by size: 2^7 / 2^7 at 2^21, then 2^8 columns (256-byte segments; 2^10 columns
// of 64 bytes at the 2^14-element stride of 2^24 took 1.64 ms for pass 1) and rows of
// 2^7 / 2^8 / 2^9 (128 / 64 / 32 rows per workgroup, 512- to 128-byte stores).
template<int LG, int N2, int NA, int C1, int G3>
constexpr ThreePassT<float> three_pass() {
    constexpr int N1 = LG - N2;
    static_assert(N1 == NA + kThreeNbLog2 && C1 << N2 == 16384 && G3 << NA == 16384, "three-pass split");
    return {LG, N2, NA, C1, 128, &launch_col<N2, C1, tw_cap<LG>()>, &launch_col<kThreeNbLog2, 128, tw_cap<N1>()>,
            &launch_row<NA, G3, LD_CONTIG, ST_STRIDED, 0, true, 0>};
}
constexpr ThreePassT<float> kThree[] = {three_pass<21, 7, 7, 128, 128>(), three_pass<22, 8, 7, 64, 128>(),
                                        three_pass<23, 8, 8, 64, 64>(), three_pass<24, 8, 9, 64, 32>()};
} // namespace
ShapeList<float> shapes_default() { return {kDefault, (int)(sizeof(kDefault) / sizeof(kDefault[0]))}; }
ThreeList<float> three_pass_tab() { return {kThree, (int)(sizeof(kThree) / sizeof(kThree[0]))}; }
} // namespace pcxd
