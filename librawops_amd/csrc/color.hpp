// Host colour math shared by the C ABI and the C++ drop-in layer.
#pragma once

#include <array>
#include <cstdint>

namespace rawops {

using Mat3 = std::array<double, 9>;

Mat3 inverse(const Mat3& m);  // throws std::invalid_argument when singular
Mat3 xyz_d50_to_working(uint32_t space);
std::array<float, 9> camera_to_working(const Mat3& camera_to_xyz, uint32_t space);
std::array<float, 9> working_to_srgb(uint32_t space);
std::array<float, 9> working_to_working(uint32_t source, uint32_t target);
// Once-rounded binary64 working-Y weights (wr, wb): Y = (g + wr(r-g)) + wb(b-g).
std::array<double, 2> working_y_weights(uint32_t space);

}  // namespace rawops
