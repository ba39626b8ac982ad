// k_xc.hip -- fused overlap-save correlation kernels, FP32 butterflies, block sizes 2 .. 16384.
#include "dispatch.hpp"

namespace pcxd {
namespace {
template<int... L>
constexpr auto xc_table(std::integer_sequence<int, L...>) {
    return std::array<XcLaunch, sizeof...(L)>{&launch_xc<L + 1>...};
}
const std::array<XcLaunch, kMaxSinglePass> kXc = xc_table(std::make_integer_sequence<int, kMaxSinglePass>{});
} // namespace

const std::array<XcLaunch, kMaxSinglePass>& xc_tab() { return kXc; }

} // namespace pcxd
