// k_f64_alt.hip -- alternative FP64 four-step shapes explored by the tuning sweeps (forward ms
// for a 2 GiB batch measured on MI355X).
#include "shapes.hpp"

namespace pcxd {
namespace {
constexpr FourStepT<double> kAlt64[] = {
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
    shape64<19, 10, 9, 8, 8>(),   // 2.44
    // 2^13-point rows, one per workgroup, and the widest column tiles (bit_reversed candidates)
    shape64<21, 13, 8, 1, 32>(), shape64<22, 13, 9, 1, 16>(), shape64<23, 13, 10, 1, 8>()};
} // namespace
ShapeList<double> shapes64_alt() { return {kAlt64, (int)(sizeof(kAlt64) / sizeof(kAlt64[0]))}; }
} // namespace pcxd
