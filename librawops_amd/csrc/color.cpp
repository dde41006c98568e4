// Colour-matrix math on the host, FP64, in the reference's operation order
// (RawEngine.cpp:116-219).  The composed matrices are cast to float32 once,
// as the reference's node constructors do (RawEngine.cpp:502-530, 603-672),
// so the device sees identical coefficients.  Built without FP contraction.
#include "color.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>

namespace rawops {
namespace {

Mat3 multiply(const Mat3& a, const Mat3& b) {
    Mat3 out{};
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col)
            for (int k = 0; k < 3; ++k) out[row * 3 + col] += a[row * 3 + k] * b[k * 3 + col];
    return out;
}

std::array<double, 3> multiply(const Mat3& a, const std::array<double, 3>& v) {
    return {a[0] * v[0] + a[1] * v[1] + a[2] * v[2],
            a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
            a[6] * v[0] + a[7] * v[1] + a[8] * v[2]};
}

std::array<double, 3> white_xyz(double x, double y) { return {x / y, 1.0, (1.0 - x - y) / y}; }

struct Primaries {
    std::array<double, 6> xy;
    double wx, wy;
};
constexpr Primaries kRomm{{0.7347, 0.2653, 0.1596, 0.8404, 0.0366, 0.0001}, 0.3457, 0.3585};
constexpr Primaries kBt2020{{0.708, 0.292, 0.170, 0.797, 0.131, 0.046}, 0.3127, 0.3290};
constexpr Primaries kSrgb{{0.640, 0.330, 0.300, 0.600, 0.150, 0.060}, 0.3127, 0.3290};

Mat3 rgb_to_xyz(const Primaries& p) {
    Mat3 basis{};
    for (int col = 0; col < 3; ++col) {
        const double x = p.xy[2 * col], y = p.xy[2 * col + 1];
        basis[col] = x / y;
        basis[3 + col] = 1.0;
        basis[6 + col] = (1.0 - x - y) / y;
    }
    const auto scales = multiply(inverse(basis), white_xyz(p.wx, p.wy));
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) basis[row * 3 + col] *= scales[col];
    return basis;
}

// Bradford D50 -> D65 adaptation.
Mat3 d50_to_d65() {
    constexpr Mat3 bradford{0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296};
    const auto d50 = multiply(bradford, white_xyz(0.3457, 0.3585));
    const auto d65 = multiply(bradford, white_xyz(0.3127, 0.3290));
    const Mat3 scale{d65[0] / d50[0], 0, 0, 0, d65[1] / d50[1], 0, 0, 0, d65[2] / d50[2]};
    return multiply(inverse(bradford), multiply(scale, bradford));
}

std::array<float, 9> to_float(const Mat3& m) {
    std::array<float, 9> out{};
    for (int i = 0; i < 9; ++i) out[i] = static_cast<float>(m[i]);
    return out;
}

}  // namespace

Mat3 inverse(const Mat3& m) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (!std::isfinite(det) || std::abs(det) < 1e-12) throw std::invalid_argument("color matrix is singular");
    const double d = 1.0 / det;
    return {(m[4] * m[8] - m[5] * m[7]) * d, (m[2] * m[7] - m[1] * m[8]) * d,
            (m[1] * m[5] - m[2] * m[4]) * d, (m[5] * m[6] - m[3] * m[8]) * d,
            (m[0] * m[8] - m[2] * m[6]) * d, (m[2] * m[3] - m[0] * m[5]) * d,
            (m[3] * m[7] - m[4] * m[6]) * d, (m[1] * m[6] - m[0] * m[7]) * d,
            (m[0] * m[4] - m[1] * m[3]) * d};
}

Mat3 xyz_d50_to_working(uint32_t space) {
    if (space == 0) return inverse(rgb_to_xyz(kRomm));
    if (space == 1) return multiply(inverse(rgb_to_xyz(kBt2020)), d50_to_d65());
    throw std::invalid_argument("unknown working space");
}

std::array<float, 9> camera_to_working(const Mat3& camera_to_xyz, uint32_t space) {
    for (double v : camera_to_xyz)
        if (!std::isfinite(v)) throw std::invalid_argument("camera-to-XYZ matrix must be finite");
    (void)inverse(camera_to_xyz);  // a collapsed camera axis cannot be recovered
    const auto composed = multiply(xyz_d50_to_working(space), camera_to_xyz);
    for (double v : composed)
        if (!std::isfinite(v) || std::abs(v) > std::numeric_limits<float>::max())
            throw std::invalid_argument("composed color matrix exceeds float32 range");
    return to_float(composed);
}

std::array<float, 9> working_to_srgb(uint32_t space) {
    const auto srgb_to_xyz = rgb_to_xyz(kSrgb);
    if (space == 0) return to_float(multiply(inverse(srgb_to_xyz), multiply(d50_to_d65(), rgb_to_xyz(kRomm))));
    if (space == 1) return to_float(multiply(inverse(srgb_to_xyz), rgb_to_xyz(kBt2020)));
    throw std::invalid_argument("unknown working space");
}

std::array<float, 9> working_to_working(uint32_t source, uint32_t target) {
    return to_float(multiply(xyz_d50_to_working(target), inverse(xyz_d50_to_working(source))));
}

std::array<double, 2> working_y_weights(uint32_t space) {
    if (space == 0) return {0.28807112822929337, 0.00008565396060525903};
    if (space == 1) return {0.26270021201126703, 0.059301716469861945};
    throw std::invalid_argument("unknown working space");
}

}  // namespace rawops
