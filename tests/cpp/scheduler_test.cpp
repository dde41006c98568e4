// TileScheduler checks.
//   scheduler_test cpu   queue semantics with a host-only probe node (no GPU
//                        calls): priority order, FIFO, pending budget,
//                        cancellation, "latest" supersession of queued and
//                        running requests, caller tokens (the reference's
//                        edit_graph_tests.cpp:1773-1911 scenarios).
//   scheduler_test gpu   device renders through the scheduler: queued RAW
//                        renders served by one batched launch equal their
//                        separate renders bit for bit; concurrent workers;
//                        TileCache clear racing an in-flight render (the
//                        cache holds its tiles in device memory).
// Exit code 0 = all checks passed; one line per check.
#include <EditGraph.hpp>
#include <TileScheduler.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <mutex>
#include <random>
#include <string>
#include <vector>

using namespace rawengine;

static int failures = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (cond) std::printf("ok   %s\n", what);                          \
        else { std::printf("FAIL %s\n", what); ++failures; }               \
    } while (0)

template <class T> static bool cancelled(std::future<T>& f) {
    try { f.get(); } catch (const RenderCancelled&) { return true; } catch (...) { return false; }
    return false;
}

// A probe that never reaches its blocking render ends the program instead of
// hanging it.
static void wait_entered(std::future<void>& f) {
    if (f.wait_for(std::chrono::seconds(30)) != std::future_status::ready) {
        std::printf("FAIL probe render never started\n");
        std::_Exit(1);
    }
}

struct ProbeState {
    std::promise<void> entered, release;
    std::shared_future<void> release_future = release.get_future().share();
    std::mutex mutex;
    std::vector<uint32_t> order;
};

// A host node: x == 0 blocks until released; records the order of renders.
class ProbeNode final : public Node {
public:
    explicit ProbeNode(std::shared_ptr<ProbeState> s) : s_(std::move(s)) {}
    Tile render(Rect b) const override {
        if (b.x == 0) {
            s_->entered.set_value();
            s_->release_future.wait();
        }
        {
            std::lock_guard lock(s_->mutex);
            s_->order.push_back(b.x);
        }
        return Tile{b, std::vector<float>(size_t(b.width) * b.height * 3, float(b.x)), output_descriptor()};
    }
    ImageDescriptor output_descriptor() const noexcept override {
        return ImageDescriptor::scene_linear(WorkingSpace::LinearProPhotoD50);
    }
private:
    std::shared_ptr<ProbeState> s_;
};

static void cpu_checks() {
    const Rect sb{0, 0, 4, 1};
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        TileScheduler sched(1, 2);
        auto first = sched.submit(node, sb, Rect{0, 0, 1, 1});
        wait_entered(entered);
        auto background = sched.submit(node, sb, Rect{1, 0, 1, 1}, RenderPriority::Background);
        auto interactive = sched.submit(node, sb, Rect{2, 0, 1, 1}, RenderPriority::Interactive);
        bool budget = false;
        try { sched.submit(node, sb, Rect{3, 0, 1, 1}); } catch (const std::length_error&) { budget = true; }
        CHECK(budget, "pending budget full -> length_error");
        st->release.set_value();
        const bool values = first.get().rgb[0] == 0 && interactive.get().rgb[0] == 2 && background.get().rgb[0] == 1;
        CHECK(values && st->order == std::vector<uint32_t>({0, 2, 1}), "queued interactive work runs before background");
        const RenderRequest typed{{3, 0, 1, 1}, 1, {}};
        CHECK(sched.submit(node, sb, typed).get().rgb[0] == 3 &&
                  sched.submit_latest("typed", node, sb, typed).get().rgb[0] == 3,
              "typed and latest requests render");
        auto unsupported = typed;
        unsupported.level.mip = 1;
        bool rejected = false;
        try { sched.submit(node, sb, unsupported); } catch (const std::invalid_argument&) { rejected = true; }
        bool empty_group = false;
        try { sched.submit_latest("", node, sb, typed); } catch (const std::invalid_argument&) { empty_group = true; }
        CHECK(rejected && empty_group, "unsupported level and empty latest group rejected at submit");
        auto token = std::make_shared<CancellationToken>();
        token->cancel();
        auto aborted = sched.submit(node, sb, Rect{3, 0, 1, 1}, RenderPriority::Normal, 256, token);
        CHECK(cancelled(aborted), "a cancelled request's future raises RenderCancelled");
    }
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        TileScheduler sched(1, 1);
        auto blocking = sched.submit(node, sb, Rect{0, 0, 1, 1});
        wait_entered(entered);
        auto obsolete = sched.submit_latest("viewport", node, sb, Rect{1, 0, 1, 1});
        auto newest = sched.submit_latest("viewport", node, sb, Rect{2, 0, 1, 1});
        const bool gone = cancelled(obsolete);
        st->release.set_value();
        CHECK(gone && blocking.get().rgb[0] == 0 && newest.get().rgb[0] == 2 &&
                  st->order == std::vector<uint32_t>({0, 2}),
              "a newer latest request replaces the queued one (budget net of the group)");
    }
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        TileScheduler sched(1, 1);
        auto blocker = sched.submit(node, sb, Rect{0, 0, 1, 1});
        wait_entered(entered);
        auto old_token = std::make_shared<CancellationToken>(), new_token = std::make_shared<CancellationToken>();
        auto old_req = sched.submit_latest("viewport", node, sb, RenderRequest{{1, 0, 1, 1}},
                                           RenderPriority::Interactive, old_token);
        auto new_req = sched.submit_latest("viewport", node, sb, Rect{2, 0, 1, 1}, RenderPriority::Interactive, 256,
                                           new_token);
        new_token->cancel();
        st->release.set_value();
        blocker.get();
        CHECK(cancelled(old_req) && cancelled(new_req) && old_token->is_cancelled() && new_token->is_cancelled() &&
                  st->order == std::vector<uint32_t>({0}),
              "supersession cancels the caller's token; a caller-cancelled latest request never renders");
    }
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        TileScheduler sched(1, 1);
        auto active = sched.submit_latest("viewport", node, sb, Rect{0, 0, 1, 1});
        wait_entered(entered);
        auto replacement = sched.submit_latest("viewport", node, sb, Rect{1, 0, 1, 1});
        st->release.set_value();
        CHECK(cancelled(active) && replacement.get().rgb[0] == 1,
              "a running latest request is cancelled by its replacement, which completes");
    }
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        std::future<Tile> dropped;
        std::future<Tile> running;
        {
            TileScheduler sched(1, 4);
            running = sched.submit(node, sb, Rect{0, 0, 1, 1});
            wait_entered(entered);
            dropped = sched.submit(node, sb, Rect{1, 0, 1, 1});
            std::thread release([&] {
                std::this_thread::sleep_for(std::chrono::milliseconds(50));
                st->release.set_value();
            });
            release.detach();
        }  // destructor: drops pending, waits for the running request
        CHECK(cancelled(dropped) && running.get().rgb[0] == 0, "destruction drops pending requests, finishes running ones");
    }
    bool bad_workers = false;
    try { TileScheduler(0, 1); } catch (const std::invalid_argument&) { bad_workers = true; }
    try { TileScheduler(65, 1); bad_workers = false; } catch (const std::invalid_argument&) {}
    CHECK(bad_workers, "worker count outside 1..64 rejected");
}

static std::vector<uint16_t> random_samples(size_t n, unsigned seed, int hi) {
    std::mt19937 rng(seed);
    std::vector<uint16_t> s(n);
    for (auto& v : s) v = uint16_t(rng() % hi);
    return s;
}

static void gpu_checks() {
    {
        auto st = std::make_shared<ProbeState>();
        auto entered = st->entered.get_future();
        auto node = std::make_shared<ProbeNode>(st);
        TileCache cache(1024);
        std::array<std::uint8_t, 32> sig{};
        sig[0] = 1;
        auto in_flight = std::async(std::launch::async, [&] { return cache.render(*node, sig, {0, 0, 1, 1}); });
        wait_entered(entered);
        cache.clear();
        st->release.set_value();
        CHECK(in_flight.get().rgb.size() == 3 && cache.stats().entries == 0,
              "TileCache: clear() keeps an in-flight render out of the cache");
    }
    RawMetadata m;
    m.width = 640; m.height = 480; m.pattern = BayerPattern::GRBG;
    m.black_levels.fill(512); m.white_levels.fill(16383);
    RawImage raw(m, random_samples(640 * 480, 11, 16383));
    GraphRecipe r;
    r.red_gain = 2.0f; r.blue_gain = 1.5f; r.exposure_stops = 0.3f;
    r.camera_color = CameraColorTransform{{0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88}, WorkingSpace::LinearProPhotoD50};
    r.output_mode = OutputMode::SrgbPreview;
    auto menon = std::make_shared<ImageGraph>(raw, r, RawDemosaicIdentity{"rawengine.menon_base", 1});
    auto bilinear = std::make_shared<ImageGraph>(raw, r);
    auto handle = [](const std::shared_ptr<ImageGraph>& g) {
        return std::shared_ptr<const Node>(g, &g->output());
    };
    // A blocking probe holds the only worker while the RAW renders queue up.
    auto st = std::make_shared<ProbeState>();
    auto entered = st->entered.get_future();
    auto probe = std::make_shared<ProbeNode>(st);
    TileScheduler sched(1, 64);
    auto blocker = sched.submit(probe, {0, 0, 4, 1}, Rect{0, 0, 1, 1}, RenderPriority::Interactive);
    wait_entered(entered);
    struct Req { std::shared_ptr<ImageGraph> g; RenderRequest r; std::future<Tile> f; };
    std::vector<Req> reqs;
    std::mt19937 rng(3);
    for (int i = 0; i < 12; ++i) {
        const bool reduced = i % 3 == 2;
        auto g = i % 2 ? menon : bilinear;
        const uint32_t mip = reduced ? 1 + i % 2 : 0;
        const uint32_t lw = (640 + (1u << mip) - 1) >> mip, lh = (480 + (1u << mip) - 1) >> mip;
        const uint32_t w = 16 + rng() % (lw - 16), h = 16 + rng() % (lh - 16);
        const Rect v{uint32_t(rng() % (lw - w + 1)), uint32_t(rng() % (lh - h + 1)), w, h};
        RenderRequest q{v, 256, reduced ? RenderLevel{mip, RenderQuality::Preview} : RenderLevel{}};
        reqs.push_back({g, q, sched.submit(handle(g), g->source_bounds(), q)});
    }
    st->release.set_value();
    blocker.get();
    bool same = true;
    for (auto& q : reqs) {
        const Tile got = q.f.get();
        const Tile want = Renderer{}.render_image(*q.g, q.r);
        same = same && got.rgb.size() == want.rgb.size() &&
               !std::memcmp(got.rgb.data(), want.rgb.data(), got.rgb.size() * sizeof(float)) &&
               got.descriptor == want.descriptor;
    }
    const auto s = sched.stats();
    std::printf("     batched launches %llu serving %llu requests\n", (unsigned long long)s.batched_launches,
                (unsigned long long)s.batched_requests);
    CHECK(same, "scheduled RAW renders (bilinear/Menon, mip 0/1/2) == separate renders (bit-exact)");
    CHECK(s.batched_launches >= 1 && s.batched_requests >= 4, "queued single-launch renders share a batched launch");
    // Several workers, several streams, a guided-filter chain among them.
    TileScheduler wide(4, 32);
    std::vector<std::future<Tile>> fs;
    std::vector<RenderRequest> rs;
    for (int i = 0; i < 16; ++i) {
        RenderRequest q{Rect{uint32_t(i * 17), uint32_t(i * 11), 200, 150}, 64};
        rs.push_back(q);
        fs.push_back(wide.submit(handle(i % 2 ? menon : bilinear), menon->source_bounds(), q,
                                 RenderPriority(i % 3)));
    }
    bool wide_same = true;
    for (int i = 0; i < 16; ++i) {
        const Tile got = fs[i].get();
        const Tile want = Renderer{}.render_image(i % 2 ? *menon : *bilinear, rs[i]);
        wide_same = wide_same && !std::memcmp(got.rgb.data(), want.rgb.data(), got.rgb.size() * sizeof(float));
    }
    CHECK(wide_same, "four workers on four streams: every request == its separate render");
}

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") cpu_checks();
    else gpu_checks();
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
