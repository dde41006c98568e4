// rawengine SpatialOps subset for the render path (reference
// SpatialOps.hpp:42-69): the working-Y guided filter node, run on the GPU
// as FP64 box moments with a fused post-filter op program.
#pragma once

#include "RawEngine.hpp"

namespace rawengine {

struct WorkingYGuidedFilterSettings {
    std::uint32_t radius = 3;    // 0..8 (reference range); requested-level pixels
    double epsilon = 0x1p-12;    // finite [2^-24, 65536]
};
void validate_working_y_guided_filter_settings(const WorkingYGuidedFilterSettings& settings);
Rect working_y_guided_filter_region(Rect output, Rect image_bounds, std::uint32_t radius);

class WorkingYGuidedFilterNode final : public DeviceNode {
public:
    WorkingYGuidedFilterNode(std::shared_ptr<const Node> input, Rect native_bounds,
                             WorkingYGuidedFilterSettings settings = {});
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region(Rect output, Rect source_bounds) const override;
    Rect input_region_level(Rect output, Rect source_bounds, RenderLevel level) const override;
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
    const WorkingYGuidedFilterSettings& settings() const noexcept { return settings_; }
    Rect native_bounds() const noexcept { return native_bounds_; }
private:
    std::shared_ptr<const Node> input_;
    Rect native_bounds_;
    WorkingYGuidedFilterSettings settings_;
    ImageDescriptor descriptor_;
};

}  // namespace rawengine
