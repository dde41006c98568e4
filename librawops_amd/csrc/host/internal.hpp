// Internals shared by the drop-in layer's translation units (RawEngine.cpp,
// EditGraph.cpp, TileScheduler.cpp): status mapping, device buffers, the
// per-thread HIP stream, cached graph nodes and the device side of TileCache.
#pragma once

#include "../../../include/rawengine/EditGraph.hpp"
#include "../../../include/rawengine/RawEngine.hpp"
#include "../../../include/rawops_hip.h"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <string>

namespace rawengine::detail {

using Signature = std::array<std::uint8_t, 32>;

[[noreturn]] void raise(rawops_status s);
inline void check(rawops_status s) {
    if (s != RAWOPS_OK) raise(s);
}
void hip(hipError_t e, const char* what);

// The HIP stream of the calling host thread: created on first use (default
// priority, not synchronising with the null stream) and handed to the next
// thread once this one exits.  Renders issued by different threads (scheduler
// workers, concurrent callers) run on different streams.
hipStream_t thread_stream();

// Device-to-host copy of `bytes` into pageable memory on `stream`, staged
// through two pinned buffers of the calling thread so that each chunk's host
// copy overlaps the next chunk's DMA; returns when dst holds the data.
void copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t stream);

// Device scratch.  With a stream, the memory comes from the device's
// stream-ordered pool (allocation and release ordered on that stream, kept in
// the pool between renders: no device-wide synchronisation per render).
struct DeviceBuffer {
    void* p = nullptr;
    hipStream_t stream = nullptr;
    bool pooled = false;
    explicit DeviceBuffer(size_t bytes) { if (bytes) hip(hipMalloc(&p, bytes), "hipMalloc"); }
    DeviceBuffer(size_t bytes, hipStream_t s);
    ~DeviceBuffer() {
        if (!p) return;
        if (pooled) (void)hipFreeAsync(p, stream);
        else (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// A tile held in HBM: interleaved float32 RGB, row-major, width x height.
struct DeviceTile {
    void* ptr = nullptr;
    std::uint32_t width = 0, height = 0;
    ImageDescriptor descriptor;
    DeviceTile(std::uint32_t w, std::uint32_t h, ImageDescriptor d);
    ~DeviceTile();
    size_t floats() const { return size_t(width) * height * 3; }
    DeviceTile(const DeviceTile&) = delete;
    DeviceTile& operator=(const DeviceTile&) = delete;
};

// TileCache's device-side interface for the chain executor.
struct CacheAccess {
    // Counts a hit or a miss like TileCache::render; on a miss `generation`
    // receives the generation to pass to store().
    static std::shared_ptr<const DeviceTile> lookup(TileCache& cache, const Signature& sig, Rect bounds,
                                                    RenderLevel level, std::uint64_t* generation);
    // Whether a tile of `floats` float32 values fits the budget at all.
    static bool admits(const TileCache& cache, size_t floats);
    // Inserts (evicting least recently used entries) unless the cache was
    // cleared since `generation` or the key is already present.
    static void store(TileCache& cache, const Signature& sig, Rect bounds, RenderLevel level,
                      std::shared_ptr<const DeviceTile> tile, std::uint64_t generation);
};

// A graph node with its tile-cache identity (EditGraph.cpp:1797-1834): every
// query forwards to the wrapped node.  Rendering goes through the device
// chain executor, which consults the cache at launch boundaries.
class CachedNode final : public DeviceNode {
public:
    CachedNode(std::shared_ptr<const Node> node, std::shared_ptr<TileCache> cache, Signature signature)
        : node_(std::move(node)), cache_(std::move(cache)), signature_(signature) {}
    bool supports_level(RenderLevel level) const noexcept override { return node_->supports_level(level); }
    RenderLevel input_level(RenderLevel level) const override { return node_->input_level(level); }
    ImageDescriptor output_descriptor() const noexcept override { return node_->output_descriptor(); }
    Rect input_region(Rect output, Rect source_bounds) const override {
        return node_->input_region(output, source_bounds);
    }
    Rect input_region_level(Rect output, Rect source_bounds, RenderLevel level) const override {
        return node_->input_region_level(output, source_bounds, level);
    }
    const Node* input_node() const noexcept override { return node_->input_node(); }
    std::optional<IccProfileIdentity> input_icc_identity() const override { return node_->input_icc_identity(); }
    std::optional<Signature> source_fingerprint() const override { return node_->source_fingerprint(); }
    std::optional<Rect> source_bounds() const override { return node_->source_bounds(); }
    std::optional<RawDemosaicIdentity> raw_demosaic_identity() const override {
        return node_->raw_demosaic_identity();
    }
    const Node& inner() const noexcept { return *node_; }
    TileCache* cache() const noexcept { return cache_.get(); }
    const Signature& signature() const noexcept { return signature_; }

private:
    std::shared_ptr<const Node> node_;
    std::shared_ptr<TileCache> cache_;
    Signature signature_;
};

// Signature of an enabled operation (EditGraph.cpp:1736-1753); an ICC output
// operation also hashes the manifest's output profile.
Signature signature_of_operation(const EditOperation& op, const Signature& upstream,
                                 const std::optional<IccProfileIdentity>& output_profile);

// A render that one batched demosaic launch can serve: a RAW chain compiled
// into a single launch with no tile-cache boundary.  `keep` owns the plan.
struct BatchPlan {
    std::shared_ptr<const void> keep;
    const rawops_plan* plan = nullptr;
    const rawops_source* source = nullptr;
};
bool batch_plan(const Node& output, BatchPlan* out);
// Cheap structural test (no compilation, no device work) for the same
// property: a RAW source under built-in point nodes only, no cache.
bool single_launch_raw_chain(const Node& output);

Rect request_bounds(Rect source, RenderLevel level);
bool compilable(const Node& n);

// Renders `output` over request.viewport into device memory (float32 RGB,
// row stride = viewport width) on `stream`, consulting graph tile caches,
// then waits for the stream and raises the chain's output-boundary errors.
void render_to_device(const Node& output, Rect source_bounds, const RenderRequest& request, float* dst,
                      hipStream_t stream);

}  // namespace rawengine::detail
