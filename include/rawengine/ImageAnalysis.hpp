// Read-only analysis of a graph output (ImageAnalysis.hpp of the reference):
// per-channel histograms computed on the GPU (rawops_histogram_rgb) over the
// device render of the request.
#pragma once

#include "RawEngine.hpp"

#include <array>
#include <cstdint>
#include <optional>
#include <vector>

namespace rawengine {

// ImageAnalysis.hpp:7-11: bins 1..65536, equally spaced in [lower, upper].
struct HistogramOptions {
    std::uint32_t bins = 256;
    double lower = 0.0;
    double upper = 1.0;
};

// ImageAnalysis.hpp:13-17.
struct ChannelHistogram {
    std::vector<std::uint64_t> counts;
    std::uint64_t underflow = 0, overflow = 0, nonfinite = 0;
    std::optional<float> minimum, maximum;   // all finite values, including outliers
};

// ImageAnalysis.hpp:19-25.
struct RgbHistogram {
    RenderRequest request;
    HistogramOptions options;
    ImageDescriptor descriptor;
    std::uint64_t pixel_count = 0;
    std::array<ChannelHistogram, 3> channels;   // R, G, B in the output domain
};

// ImageAnalysis.hpp:27-36: bins are half-open except the last, which includes
// upper; signed values and headroom are tallied, not clipped; NaN and
// infinities count as nonfinite and stay out of the extrema.  A chain the
// library compiles is rendered on the device band by band and analysed
// there; a foreign output node's tiles are checked (bounds, storage,
// descriptor), uploaded and analysed on the device.  Cancellation throws.
RgbHistogram histogram_rgb(const Node& output, Rect source_bounds, const RenderRequest& request,
                           const HistogramOptions& options = {}, const CancellationToken* cancellation = nullptr);
RgbHistogram histogram_rgb(const ImageGraph& graph, const RenderRequest& request,
                           const HistogramOptions& options = {}, const CancellationToken* cancellation = nullptr);
// ImageAnalysis.hpp:38-40: a manifest graph's output over its output bounds.
RgbHistogram histogram_rgb(const ExecutableEditGraph& graph, const RenderRequest& request,
                           const HistogramOptions& options = {}, const CancellationToken* cancellation = nullptr);

}  // namespace rawengine
