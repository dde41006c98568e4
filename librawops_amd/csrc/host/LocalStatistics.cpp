// Local statistics (SpatialOps.cpp:72-146): tiles rendered with their halo on
// the device, moments by rawops_local_statistics, copied back per tile for the
// callback.
#include "../../../include/rawengine/EditGraph.hpp"
#include "../../../include/rawengine/SpatialOps.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <limits>

namespace rawengine {

namespace {

void addressable(Rect r) {
    if (!r.width || !r.height || uint64_t(r.x) + r.width > (uint64_t{1} << 32) ||
        uint64_t(r.y) + r.height > (uint64_t{1} << 32))
        throw std::invalid_argument("spatial operation requires nonempty addressable bounds");
}

rawops_rect c(Rect r) { return {r.x, r.y, r.width, r.height}; }

// Moments of a device tile covering `needed` into a host result.
struct Workspace {
    detail::DeviceBuffer mean, var, count;
    explicit Workspace(size_t values)
        : mean(values * sizeof(double)), var(values * sizeof(double)), count(values * sizeof(uint32_t)) {}
};

LocalStatisticsTile moments(const float* dev_in, Rect needed, Rect output, Rect image, uint32_t radius,
                            const ImageDescriptor& d, Workspace& ws, hipStream_t stream) {
    const size_t n = size_t(output.width) * output.height * 3;
    LocalStatisticsTile t{output, d, radius, std::vector<double>(n), std::vector<double>(n),
                          std::vector<std::uint32_t>(n)};
    detail::check(rawops_local_statistics(dev_in, c(needed), needed.width, c(output), c(image), radius,
                                          static_cast<double*>(ws.mean.p), static_cast<double*>(ws.var.p),
                                          static_cast<uint32_t*>(ws.count.p), stream));
    detail::hip(hipMemcpyAsync(t.mean.data(), ws.mean.p, n * sizeof(double), hipMemcpyDeviceToHost, stream), "copy");
    detail::hip(hipMemcpyAsync(t.variance.data(), ws.var.p, n * sizeof(double), hipMemcpyDeviceToHost, stream), "copy");
    detail::hip(hipMemcpyAsync(t.finite_count.data(), ws.count.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                "copy");
    detail::hip(hipStreamSynchronize(stream), "stream synchronize");
    return t;
}

void check_tile(const Tile& t, Rect needed, const ImageDescriptor& d) {
    if (t.bounds.x != needed.x || t.bounds.y != needed.y || t.bounds.width != needed.width ||
        t.bounds.height != needed.height || t.rgb.size() != size_t(needed.width) * needed.height * 3 ||
        t.descriptor != d || d.format != PixelFormat::RGBFloat32)
        throw std::invalid_argument("spatial input tile bounds, storage or descriptor differ");
}

}  // namespace

// SpatialOps.cpp:72-75.
Rect local_statistics_region(Rect output, Rect image, std::uint32_t radius) {
    if (radius > 8) throw std::invalid_argument("local statistics radius must be in 0..8");
    addressable(output);
    addressable(image);
    if (output.x < image.x || output.y < image.y ||
        uint64_t(output.x) + output.width > uint64_t(image.x) + image.width ||
        uint64_t(output.y) + output.height > uint64_t(image.y) + image.height)
        throw std::out_of_range("spatial viewport is outside image bounds");
    const auto left = std::max<int64_t>(image.x, int64_t(output.x) - radius);
    const auto top = std::max<int64_t>(image.y, int64_t(output.y) - radius);
    const auto right = std::min<uint64_t>(uint64_t(image.x) + image.width, uint64_t(output.x) + output.width + radius);
    const auto bottom = std::min<uint64_t>(uint64_t(image.y) + image.height, uint64_t(output.y) + output.height + radius);
    return {uint32_t(left), uint32_t(top), uint32_t(right - left), uint32_t(bottom - top)};
}

// SpatialOps.cpp:77-121 on a host tile: uploaded, reduced on the device.
LocalStatisticsTile local_statistics_rgb(const Tile& input, Rect output, Rect image, std::uint32_t radius) {
    const Rect needed = local_statistics_region(output, image, radius);
    check_tile(input, needed, input.descriptor);
    const hipStream_t stream = detail::thread_stream();
    detail::DeviceBuffer in(input.rgb.size() * sizeof(float));
    detail::hip(hipMemcpyAsync(in.p, input.rgb.data(), input.rgb.size() * sizeof(float), hipMemcpyHostToDevice, stream),
                "copy");
    Workspace ws(size_t(output.width) * output.height * 3);
    return moments(static_cast<const float*>(in.p), needed, output, image, radius, input.descriptor, ws, stream);
}

// SpatialOps.cpp:123-146.
void analyze_local_rgb(const Node& output, Rect native_bounds, const RenderRequest& request,
                       const LocalStatisticsCallback& callback, std::uint32_t radius,
                       const CancellationToken* cancellation) {
    addressable(native_bounds);
    const Rect image = detail::request_bounds(native_bounds, request.level);
    (void)local_statistics_region(request.viewport, image, radius);
    if (!callback || !request.tile_size || !output.supports_level(request.level))
        throw std::invalid_argument("local analysis needs a callback, positive tile size and supported level");
    auto check = [&] {
        if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
    };
    check();
    const hipStream_t stream = detail::thread_stream();
    const bool device = detail::compilable(output);
    const ImageDescriptor d = output.output_descriptor();
    const uint32_t tw = std::min(request.tile_size, request.viewport.width);
    const uint32_t th = std::min(request.tile_size, request.viewport.height);
    const size_t in_floats = size_t(tw + 2 * radius) * (th + 2 * radius) * 3;
    detail::DeviceBuffer in(in_floats * sizeof(float));
    Workspace ws(size_t(tw) * th * 3);
    const uint64_t right = uint64_t(request.viewport.x) + request.viewport.width;
    const uint64_t bottom = uint64_t(request.viewport.y) + request.viewport.height;
    for (uint64_t y = request.viewport.y; y < bottom; y += request.tile_size)
        for (uint64_t x = request.viewport.x; x < right; x += request.tile_size) {
            check();
            const Rect rect{uint32_t(x), uint32_t(y), uint32_t(std::min<uint64_t>(request.tile_size, right - x)),
                            uint32_t(std::min<uint64_t>(request.tile_size, bottom - y))};
            const Rect needed = local_statistics_region(rect, image, radius);
            if (device) {
                detail::render_to_device(output, native_bounds, RenderRequest{needed, request.tile_size, request.level},
                                         static_cast<float*>(in.p), stream);
            } else {
                const Tile t = output.render_level(needed, request.level);
                check();
                check_tile(t, needed, d);
                detail::hip(hipMemcpyAsync(in.p, t.rgb.data(), t.rgb.size() * sizeof(float), hipMemcpyHostToDevice,
                                           stream), "copy");
            }
            check();
            const auto stats = moments(static_cast<const float*>(in.p), needed, rect, image, radius, d, ws, stream);
            check();
            callback(stats);
            check();
        }
}

void analyze_local_rgb(const ImageGraph& graph, const RenderRequest& request, const LocalStatisticsCallback& callback,
                       std::uint32_t radius, const CancellationToken* cancellation) {
    analyze_local_rgb(graph.output(), graph.source_bounds(), request, callback, radius, cancellation);
}

void analyze_local_rgb(const ExecutableEditGraph& graph, const RenderRequest& request,
                       const LocalStatisticsCallback& callback, std::uint32_t radius,
                       const CancellationToken* cancellation) {
    analyze_local_rgb(graph.output(), graph.output_bounds(), request, callback, radius, cancellation);
}

}  // namespace rawengine
