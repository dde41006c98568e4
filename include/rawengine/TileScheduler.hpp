// Asynchronous render requests, MI355X build (TileScheduler.hpp:1-86 of the
// reference, RGB outputs).  Queued requests run in priority order, FIFO within
// one priority; running requests are not preempted; the pending budget, the
// "latest" supersession groups and shared cancellation tokens behave as the
// reference's (TileScheduler.cpp:19-246).
//
// Execution: each worker thread owns a HIP stream, so up to `workers`
// requests are in flight on the GPU at once.  When a worker takes a request
// and at least three more requests of the same priority are queued whose
// chains compile into a single demosaic launch (a RAW source with point
// operations, no tile cache), it takes up to 64 of them, in queue order, and
// renders them with one batched launch (one workgroup per (request, tile)
// item, include/rawops_hip.h rawops_batch_*): results equal separate renders
// bit for bit.
#pragma once

#include "RawEngine.hpp"

#include <cstddef>
#include <cstdint>
#include <future>
#include <memory>
#include <string>

namespace rawengine {

enum class RenderPriority { Background = 0, Normal = 1, Interactive = 2 };

class TileScheduler final {
public:
    // 1..64 workers and a positive pending budget (std::invalid_argument).
    TileScheduler(std::size_t workers, std::size_t max_pending);
    // Drops pending requests (their futures raise RenderCancelled) and waits
    // for running ones.
    ~TileScheduler();
    TileScheduler(const TileScheduler&) = delete;
    TileScheduler& operator=(const TileScheduler&) = delete;

    // std::length_error when the pending queue is full.  The token is checked
    // before the render, at its tile boundaries and after it: a cancelled
    // request's future raises RenderCancelled.
    std::future<Tile> submit(std::shared_ptr<const Node> output, Rect source_bounds, RenderRequest request,
                             RenderPriority priority = RenderPriority::Normal,
                             std::shared_ptr<CancellationToken> cancellation = nullptr);
    std::future<Tile> submit(std::shared_ptr<const Node> output, Rect source_bounds, Rect viewport,
                             RenderPriority priority = RenderPriority::Normal, std::uint32_t tile_size = 256,
                             std::shared_ptr<CancellationToken> cancellation = nullptr);

    // Only the newest request of a nonempty group stays actionable: older
    // queued requests of the group fail with RenderCancelled and a running one
    // is cancelled at its next tile boundary.
    std::future<Tile> submit_latest(std::string group, std::shared_ptr<const Node> output, Rect source_bounds,
                                    RenderRequest request, RenderPriority priority = RenderPriority::Interactive,
                                    std::shared_ptr<CancellationToken> cancellation = nullptr);
    std::future<Tile> submit_latest(std::string group, std::shared_ptr<const Node> output, Rect source_bounds,
                                    Rect viewport, RenderPriority priority = RenderPriority::Interactive,
                                    std::uint32_t tile_size = 256,
                                    std::shared_ptr<CancellationToken> cancellation = nullptr);

    // Extension: counters of the device launches made so far.
    struct Stats {
        std::uint64_t completed = 0;          // requests settled with a tile
        std::uint64_t batched_launches = 0;   // batched device launches
        std::uint64_t batched_requests = 0;   // requests served by them
    };
    Stats stats() const;

private:
    std::future<Tile> submit_request(std::string group, std::shared_ptr<const Node> output, Rect source_bounds,
                                     RenderRequest request, RenderPriority priority,
                                     std::shared_ptr<CancellationToken> cancellation);
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace rawengine
