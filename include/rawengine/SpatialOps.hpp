// rawengine SpatialOps subset for the render path (reference
// SpatialOps.hpp:7-69): local statistics and the working-Y guided filter
// node, run on the GPU (FP64 moments; the filter with a fused post-filter op
// program).
#pragma once

#include "RawEngine.hpp"

namespace rawengine {

class ExecutableEditGraph;

// SpatialOps.hpp:7-13: local moments of a rendered tile, computed on the GPU
// in FP64 in the reference's operation order (equal to its values bit for bit).
struct LocalStatisticsTile {
    Rect bounds;
    ImageDescriptor descriptor;
    std::uint32_t radius = 3;                  // 0..8, requested-level pixels
    std::vector<double> mean, variance;        // interleaved RGB; population variance
    std::vector<std::uint32_t> finite_count;   // non-finite samples excluded per channel
};

// SpatialOps.hpp:15-34.  Windows clip at the true image edge; the input tile
// must cover local_statistics_region(output, image, radius).  Empty finite
// windows have NaN moments.
Rect local_statistics_region(Rect output, Rect image_bounds, std::uint32_t radius);
LocalStatisticsTile local_statistics_rgb(const Tile& input, Rect output, Rect image_bounds, std::uint32_t radius = 3);
using LocalStatisticsCallback = std::function<void(const LocalStatisticsTile&)>;
// Streams one statistics tile per render tile (each rendered with its halo on
// the device); the callback's storage is temporary.  Cancellation and errors
// stop delivery; earlier callbacks remain.
void analyze_local_rgb(const Node& output, Rect native_bounds, const RenderRequest& request,
                       const LocalStatisticsCallback& callback, std::uint32_t radius = 3,
                       const CancellationToken* cancellation = nullptr);
void analyze_local_rgb(const ImageGraph& graph, const RenderRequest& request, const LocalStatisticsCallback& callback,
                       std::uint32_t radius = 3, const CancellationToken* cancellation = nullptr);
void analyze_local_rgb(const ExecutableEditGraph& graph, const RenderRequest& request,
                       const LocalStatisticsCallback& callback, std::uint32_t radius = 3,
                       const CancellationToken* cancellation = nullptr);

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
