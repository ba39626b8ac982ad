// k_shapes_alt.hip -- alternative FP32 four-step shapes explored by the tuning sweeps and the
// uneven-batch autotuner.
#include "shapes.hpp"

namespace pcxd {
namespace {
constexpr FourStep kAlt[] = {
    // N = 2^16
    shape<16, 8, 8, 32, 64>(), shape<16, 8, 8, 16, 64>(), shape<16, 8, 8, 64, 64>(), shape<16, 8, 8, 32, 32>(),
    shape<16, 8, 8, 16, 32>(), shape<16, 10, 6, 16, 64>(), shape<16, 7, 9, 64, 32>(), shape<16, 6, 10, 64, 16>(),
    shape<16, 9, 7, 32, 32>(), shape<16, 10, 6, 16, 32>(), shape<16, 9, 7, 32, 64>(),
    // 2^20 permuted orders (the default there): 4096-point rows, 64-column tiles
    shape<20, 12, 8, 4, 64>(),
    // N = 2^15, 2^19 (the second: 2^19's former default, 16-column tiles)
    shape<15, 8, 7, 32, 64>(), shape<15, 8, 7, 32, 128>(), shape<19, 10, 9, 16, 32>(), shape<19, 9, 10, 32, 16>(),
    // the largest sizes
    shape<21, 12, 9, 4, 32>(), shape<21, 13, 8, 2, 64>(), shape<21, 14, 7, 1, 128>(), shape<22, 13, 9, 2, 32>(),
    shape<22, 14, 8, 1, 64>(), shape<23, 13, 10, 2, 16>(), shape<23, 14, 9, 1, 32>(), shape<24, 14, 10, 1, 16>(),
    shape<24, 13, 11, 2, 8>(),
};
} // namespace
ShapeList<float> shapes_alt() { return {kAlt, (int)(sizeof(kAlt) / sizeof(kAlt[0]))}; }
} // namespace pcxd
