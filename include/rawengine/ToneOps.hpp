// rawengine ToneOps subset for the render path (reference ToneOps.hpp:7-50):
// ExtendedCurvesNode, master-then-channel FP64 curves fused into the device
// op program.
#pragma once

#include "RawEngine.hpp"

namespace rawengine {

struct CurvePoint { double x = 0, y = 0; };
enum class CurveInterpolation { Linear, ShapePreservingCubic };
struct CurveKnots { std::vector<CurvePoint> knots{{0, 0}, {1, 1}}; };
struct ExtendedCurvesSettings {
    CurveKnots master;
    std::array<CurveKnots, 3> channels;
    CurveInterpolation interpolation = CurveInterpolation::Linear;
};
void validate_extended_curves_settings(const ExtendedCurvesSettings& settings);

class ExtendedCurvesNode final : public DeviceNode {
public:
    ExtendedCurvesNode(std::shared_ptr<const Node> input, ExtendedCurvesSettings settings = {});
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect bounds, RenderLevel level) const override;
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return input_->output_descriptor(); }
    const ExtendedCurvesSettings& settings() const noexcept { return settings_; }
private:
    std::shared_ptr<const Node> input_;
    ExtendedCurvesSettings settings_;
};

}  // namespace rawengine
