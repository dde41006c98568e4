// Priority queue of render requests served by worker threads, each with its
// own HIP stream (TileScheduler.cpp:19-246 of the reference for the queue,
// budget, supersession and cancellation rules).  Runs of at least four
// queued single-launch RAW renders of one priority become one batched launch.
#include "../../../include/rawengine/TileScheduler.hpp"
#include "internal.hpp"

#include <algorithm>
#include <condition_variable>
#include <exception>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

namespace rawengine {

namespace {
constexpr size_t kBatchMin = 4;   // a batch costs a graph capture: worth it from four requests
constexpr size_t kBatchMax = 64;
constexpr uint32_t kBatchShape = 0;  // 64x32 tile, the shape for small viewports
}  // namespace

struct TileScheduler::Impl {
    struct Job {
        std::shared_ptr<const Node> output;
        Rect source_bounds;
        RenderRequest request;
        RenderPriority priority = RenderPriority::Normal;
        std::shared_ptr<CancellationToken> cancellation;
        std::string group;
        std::uint64_t sequence = 0;
        bool batchable = false;
        std::promise<Tile> result;
    };

    Impl(size_t workers, size_t max_pending) : max_pending(max_pending) {
        if (!workers || workers > 64 || !max_pending)
            throw std::invalid_argument("scheduler needs 1-64 workers and a pending budget");
        try {
            for (size_t i = 0; i < workers; ++i) threads.emplace_back([this] { work(); });
        } catch (...) {
            stop();
            throw;
        }
    }

    ~Impl() { stop(); }

    void stop() {
        {
            std::lock_guard lock(mutex);
            stopping = true;
            for (auto& job : pending) job->result.set_exception(std::make_exception_ptr(RenderCancelled()));
            pending.clear();
            latest.clear();
        }
        ready.notify_all();
        for (auto& t : threads)
            if (t.joinable()) t.join();
    }

    // Highest priority first, then submission order.
    static bool before(const Job& a, const Job& b) {
        if (a.priority != b.priority) return int(a.priority) > int(b.priority);
        return a.sequence < b.sequence;
    }

    // Under the lock: the next job and, when enough of its kind are queued,
    // its batch companions (same priority, single-launch RAW, queue order).
    std::vector<std::shared_ptr<Job>> take() {
        auto best = std::min_element(pending.begin(), pending.end(),
                                     [](const auto& a, const auto& b) { return before(*a, *b); });
        std::vector<std::shared_ptr<Job>> out{*best};
        pending.erase(best);
        const Job& lead = *out.front();
        if (!lead.batchable) return out;
        std::vector<std::shared_ptr<Job>> mates;
        for (const auto& j : pending)
            if (j->batchable && j->priority == lead.priority) mates.push_back(j);
        if (mates.size() + 1 < kBatchMin) return out;
        std::sort(mates.begin(), mates.end(), [](const auto& a, const auto& b) { return a->sequence < b->sequence; });
        if (mates.size() > kBatchMax - 1) mates.resize(kBatchMax - 1);
        for (const auto& m : mates) pending.erase(std::find(pending.begin(), pending.end(), m));
        out.insert(out.end(), mates.begin(), mates.end());
        return out;
    }

    void work() {
        for (;;) {
            std::vector<std::shared_ptr<Job>> jobs;
            {
                std::unique_lock lock(mutex);
                ready.wait(lock, [&] { return stopping || !pending.empty(); });
                if (stopping && pending.empty()) return;
                jobs = take();
            }
            if (jobs.size() == 1) run_one(jobs.front());
            else run_batch(jobs);
        }
    }

    // Settles a job under the lock and retires its supersession group entry.
    void settle(const std::shared_ptr<Job>& job, std::optional<Tile> tile, std::exception_ptr failure) {
        std::lock_guard lock(mutex);
        if (!failure && job->cancellation && job->cancellation->is_cancelled())
            failure = std::make_exception_ptr(RenderCancelled());
        if (failure) {
            job->result.set_exception(failure);
        } else {
            job->result.set_value(std::move(*tile));
            ++stats.completed;
        }
        if (!job->group.empty()) {
            auto found = latest.find(job->group);
            if (found != latest.end() && found->second.lock() == job->cancellation) latest.erase(found);
        }
    }

    void run_one(const std::shared_ptr<Job>& job) {
        std::optional<Tile> tile;
        std::exception_ptr failure;
        try {
            if (job->cancellation && job->cancellation->is_cancelled()) throw RenderCancelled();
            tile = Renderer{}.render_image(*job->output, job->source_bounds, job->request, job->cancellation.get());
        } catch (...) {
            failure = std::current_exception();
        }
        settle(job, std::move(tile), failure);
    }

    // One batched device launch for the live jobs; any job the batch cannot
    // take (cancelled, empty, failing to compile) or a batch that reports a
    // non-finite output falls back to its own render, so every error belongs
    // to the request that caused it.
    void run_batch(std::vector<std::shared_ptr<Job>>& jobs) {
        std::vector<std::shared_ptr<Job>> live;
        std::vector<detail::BatchPlan> plans;
        for (auto& job : jobs) {
            detail::BatchPlan plan;
            const Rect v = job->request.viewport;
            if ((job->cancellation && job->cancellation->is_cancelled()) || !v.width || !v.height) {
                run_one(job);
                continue;
            }
            bool ok = false;
            try {
                ok = detail::batch_plan(*job->output, &plan);
            } catch (const std::exception&) {
                ok = false;
            }
            if (!ok) {
                run_one(job);
                continue;
            }
            live.push_back(job);
            plans.push_back(std::move(plan));
        }
        if (live.size() < 2) {
            for (auto& job : live) run_one(job);
            return;
        }
        std::vector<size_t> offset(live.size() + 1, 0);
        for (size_t i = 0; i < live.size(); ++i) {
            const Rect v = live[i]->request.viewport;
            offset[i + 1] = offset[i] + size_t(v.width) * v.height * 3;
        }
        std::vector<std::optional<Tile>> tiles(live.size());
        bool fallback = false;
        try {
            const hipStream_t stream = detail::thread_stream();
            detail::DeviceBuffer out(offset.back() * sizeof(float));
            std::vector<rawops_batch_request> reqs(live.size());
            for (size_t i = 0; i < live.size(); ++i) {
                const auto& r = live[i]->request;
                reqs[i] = {};
                reqs[i].plan = plans[i].plan;
                reqs[i].source = plans[i].source;
                reqs[i].viewport = {r.viewport.x, r.viewport.y, r.viewport.width, r.viewport.height};
                reqs[i].mip = r.level.mip;
                reqs[i].quality = uint32_t(r.level.quality);
                reqs[i].linear_ops = std::numeric_limits<uint32_t>::max();
                reqs[i].dst_rgb = static_cast<float*>(out.p) + offset[i];
            }
            rawops_batch* batch = nullptr;
            detail::check(rawops_batch_create(reqs.data(), reqs.size(), kBatchShape, stream, &batch));
            struct Destroy { rawops_batch* b; ~Destroy() { rawops_batch_destroy(b); } } destroy{batch};
            detail::check(rawops_batch_run(batch, stream));
            if (rawops_batch_check(batch, stream) != RAWOPS_OK) {
                fallback = true;
            } else {
                for (size_t i = 0; i < live.size(); ++i) {
                    const auto& r = live[i]->request;
                    Tile t{r.viewport, std::vector<float>(offset[i + 1] - offset[i]), live[i]->output->output_descriptor()};
                    detail::hip(hipMemcpyAsync(t.rgb.data(), static_cast<float*>(out.p) + offset[i],
                                               t.rgb.size() * sizeof(float), hipMemcpyDeviceToHost, stream),
                                "hipMemcpyAsync(batch output)");
                    tiles[i] = std::move(t);
                }
                detail::hip(hipStreamSynchronize(stream), "stream synchronize");
            }
        } catch (...) {
            fallback = true;
        }
        if (fallback) {
            for (auto& job : live) run_one(job);
            return;
        }
        {
            std::lock_guard lock(mutex);
            ++stats.batched_launches;
            stats.batched_requests += live.size();
        }
        for (size_t i = 0; i < live.size(); ++i) settle(live[i], std::move(tiles[i]), nullptr);
    }

    // TileScheduler.cpp:122-152: budget net of the group's queued requests,
    // supersession of queued and running group members.
    void enqueue(std::shared_ptr<Job> job) {
        {
            std::lock_guard lock(mutex);
            size_t removable = 0;
            if (!job->group.empty())
                for (const auto& q : pending) removable += q->group == job->group;
            if (pending.size() - removable >= max_pending)
                throw std::length_error("scheduler pending request budget is full");
            if (!job->group.empty()) {
                if (auto found = latest.find(job->group); found != latest.end())
                    if (auto previous = found->second.lock()) previous->cancel();
                std::vector<std::shared_ptr<Job>> kept;
                for (auto& q : pending) {
                    if (q->group == job->group) {
                        q->cancellation->cancel();
                        q->result.set_exception(std::make_exception_ptr(RenderCancelled()));
                    } else {
                        kept.push_back(std::move(q));
                    }
                }
                pending = std::move(kept);
                latest[job->group] = job->cancellation;
            }
            job->sequence = next_sequence++;
            pending.push_back(std::move(job));
        }
        ready.notify_one();
    }

    const size_t max_pending;
    mutable std::mutex mutex;
    std::condition_variable ready;
    std::vector<std::shared_ptr<Job>> pending;
    std::map<std::string, std::weak_ptr<CancellationToken>> latest;
    std::vector<std::thread> threads;
    std::uint64_t next_sequence = 0;
    bool stopping = false;
    Stats stats;
};

TileScheduler::TileScheduler(std::size_t workers, std::size_t max_pending)
    : impl_(std::make_unique<Impl>(workers, max_pending)) {}

TileScheduler::~TileScheduler() = default;

TileScheduler::Stats TileScheduler::stats() const {
    std::lock_guard lock(impl_->mutex);
    return impl_->stats;
}

std::future<Tile> TileScheduler::submit(std::shared_ptr<const Node> output, Rect source_bounds, RenderRequest request,
                                        RenderPriority priority, std::shared_ptr<CancellationToken> cancellation) {
    return submit_request({}, std::move(output), source_bounds, request, priority, std::move(cancellation));
}

std::future<Tile> TileScheduler::submit(std::shared_ptr<const Node> output, Rect source_bounds, Rect viewport,
                                        RenderPriority priority, std::uint32_t tile_size,
                                        std::shared_ptr<CancellationToken> cancellation) {
    return submit(std::move(output), source_bounds, RenderRequest{viewport, tile_size, {}}, priority,
                  std::move(cancellation));
}

std::future<Tile> TileScheduler::submit_latest(std::string group, std::shared_ptr<const Node> output,
                                               Rect source_bounds, RenderRequest request, RenderPriority priority,
                                               std::shared_ptr<CancellationToken> cancellation) {
    if (group.empty()) throw std::invalid_argument("latest request group is empty");
    if (!cancellation) cancellation = std::make_shared<CancellationToken>();
    return submit_request(std::move(group), std::move(output), source_bounds, request, priority,
                          std::move(cancellation));
}

std::future<Tile> TileScheduler::submit_latest(std::string group, std::shared_ptr<const Node> output,
                                               Rect source_bounds, Rect viewport, RenderPriority priority,
                                               std::uint32_t tile_size, std::shared_ptr<CancellationToken> cancellation) {
    return submit_latest(std::move(group), std::move(output), source_bounds, RenderRequest{viewport, tile_size, {}},
                         priority, std::move(cancellation));
}

// TileScheduler.cpp:205-246: the request is validated before it is queued.
std::future<Tile> TileScheduler::submit_request(std::string group, std::shared_ptr<const Node> output,
                                                Rect source_bounds, RenderRequest request, RenderPriority priority,
                                                std::shared_ptr<CancellationToken> cancellation) {
    const Rect v = request.viewport;
    const RenderLevel l = request.level;
    const bool native = l.mip == 0 && (l.quality == RenderQuality::Final || l.quality == RenderQuality::Preview);
    const bool reduced = l.mip >= 1 && l.mip <= 2 && l.quality == RenderQuality::Preview;
    const uint32_t scale = reduced ? 1u << l.mip : 1u;
    const Rect bounds = reduced ? Rect{0, 0, source_bounds.width / scale + (source_bounds.width % scale != 0),
                                       source_bounds.height / scale + (source_bounds.height % scale != 0)}
                                : source_bounds;
    const uint64_t limit = uint64_t(std::numeric_limits<uint32_t>::max()) + 1;
    if (!output || !request.tile_size || (!native && !reduced) || !output->supports_level(l) ||
        !source_bounds.width || !source_bounds.height || int(priority) < int(RenderPriority::Background) ||
        int(priority) > int(RenderPriority::Interactive) || uint64_t(source_bounds.x) + source_bounds.width > limit ||
        uint64_t(source_bounds.y) + source_bounds.height > limit || v.x < bounds.x || v.y < bounds.y ||
        uint64_t(v.x) + v.width > uint64_t(bounds.x) + bounds.width ||
        uint64_t(v.y) + v.height > uint64_t(bounds.y) + bounds.height)
        throw std::invalid_argument("invalid scheduled render request");
    auto job = std::make_shared<Impl::Job>();
    job->batchable = detail::single_launch_raw_chain(*output);
    job->output = std::move(output);
    job->source_bounds = source_bounds;
    job->request = request;
    job->priority = priority;
    job->cancellation = std::move(cancellation);
    job->group = std::move(group);
    auto result = job->result.get_future();
    impl_->enqueue(std::move(job));
    return result;
}

}  // namespace rawengine
