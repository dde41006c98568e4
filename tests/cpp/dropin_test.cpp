// Drop-in API test: the rawengine:: C++ layer (include/rawengine/*.hpp) on the
// GPU, checked against the CPU oracle (oracle/liboracle.so) on the same
// seeded inputs.  Exit code 0 = all checks passed; prints one line per check.
#include <RawEngine.hpp>
#include <SpatialOps.hpp>
#include <ToneOps.hpp>
#include <ImageAnalysis.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../oracle/rawops_oracle.h"

using namespace rawengine;

static int failures = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (cond) std::printf("ok   %s\n", what);                          \
        else { std::printf("FAIL %s\n", what); ++failures; }               \
    } while (0)

template <class F> static bool throws_as(F f, int kind) {
    try { f(); } catch (const RenderCancelled&) { return kind == 4; }
    catch (const std::out_of_range&) { return kind == 1; }
    catch (const std::invalid_argument&) { return kind == 2; }
    catch (const std::domain_error&) { return kind == 3; }
    catch (...) { return false; }
    return false;
}

static std::vector<uint16_t> random_samples(size_t n, unsigned seed, int hi) {
    std::mt19937 rng(seed);
    std::vector<uint16_t> s(n);
    for (auto& v : s) v = uint16_t(rng() % hi);
    return s;
}

static ro_bayer oracle_meta(const RawMetadata& m) {
    ro_bayer b{};
    b.width = m.width; b.height = m.height; b.row_stride = m.row_stride_samples ? m.row_stride_samples : m.width;
    b.pattern = uint32_t(m.pattern); b.phase_x = m.cfa_phase_x; b.phase_y = m.cfa_phase_y;
    Rect a = m.active_area;
    if (!a.width) a = {0, 0, m.width, m.height};
    b.area = {a.x, a.y, a.width, a.height};
    for (int i = 0; i < 4; ++i) { b.black[i] = m.black_levels[i]; b.white[i] = m.white_levels[i]; }
    return b;
}

static bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
    return a.size() == b.size() && !std::memcmp(a.data(), b.data(), a.size() * sizeof(float));
}

// A node type the library does not know: constant camera-linear pixels.
class ConstantCamera final : public Node {
public:
    Tile render(Rect b) const override {
        Tile t{b, std::vector<float>(size_t(b.width) * b.height * 3)};
        for (size_t i = 0; i < t.rgb.size(); i += 3) { t.rgb[i] = 0.5f; t.rgb[i + 1] = 0.25f; t.rgb[i + 2] = 0.1f; }
        return t;
    }
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::camera_linear(); }
};

// A foreign node for histogram_rgb: R = (x mod 9) / 4 - 1/2, G = (y mod 9) / 4 - 1/2, B = NaN;
// optionally a malformed tile or a cancellation from inside render().
class ResidueProbe final : public Node {
public:
    int fault = 0;                         // 1 bounds, 2 storage, 3 descriptor
    CancellationToken* cancel_on_render = nullptr;
    mutable unsigned calls = 0;
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::camera_linear(); }
    Tile render(Rect r) const override {
        ++calls;
        Tile t{r, {}, output_descriptor()};
        for (uint32_t y = r.y; y < r.y + r.height; ++y)
            for (uint32_t x = r.x; x < r.x + r.width; ++x) {
                t.rgb.push_back(float(x % 9) * 0.25f - 0.5f);
                t.rgb.push_back(float(y % 9) * 0.25f - 0.5f);
                t.rgb.push_back(std::nanf(""));
            }
        if (fault == 1) ++t.bounds.x;
        if (fault == 2) t.rgb.pop_back();
        if (fault == 3) t.descriptor = ImageDescriptor::srgb_output();
        if (cancel_on_render) cancel_on_render->cancel();
        return t;
    }
};

static void histogram_checks() {
    ResidueProbe probe;
    const Rect bounds{7, 5, 31, 23};
    const RenderRequest request{Rect{9, 7, 19, 13}, 4};
    const auto h = histogram_rgb(probe, bounds, request, HistogramOptions{4, 0.0, 1.0});
    bool edges = h.pixel_count == 247;
    for (int c = 0; c < 2; ++c) {
        uint64_t want[4] = {0, 0, 0, 0}, below = 0, above = 0;
        for (uint32_t y = 7; y < 20; ++y)
            for (uint32_t x = 9; x < 28; ++x) {
                const uint32_t r = (c == 0 ? x : y) % 9;   // value r / 4 - 1/2
                if (r < 2) ++below;
                else if (r > 6) ++above;
                else ++want[r == 6 ? 3 : r - 2];           // 1.0 joins the closed last bin
            }
        edges = edges && h.channels[c].counts == std::vector<uint64_t>(want, want + 4) &&
                h.channels[c].underflow == below && h.channels[c].overflow == above &&
                h.channels[c].nonfinite == 0 && h.channels[c].minimum == -0.5f && h.channels[c].maximum == 1.5f;
    }
    edges = edges && h.channels[2].nonfinite == 247 && !h.channels[2].minimum && !h.channels[2].maximum;
    CHECK(edges, "histogram_rgb: bin edges, outliers, nonfinite, extrema over a foreign node's tiles");
    bool partition = true;
    // tile sizes far beyond the viewport stage min(tile_size, viewport) per side
    for (uint32_t ts : {1u, 3u, 7u, 256u, 65536u, 0x80000000u}) {
        const auto o = histogram_rgb(probe, bounds, RenderRequest{request.viewport, ts}, HistogramOptions{4, 0.0, 1.0});
        for (int c = 0; c < 3; ++c)
            partition = partition && o.channels[c].counts == h.channels[c].counts &&
                        o.channels[c].underflow == h.channels[c].underflow &&
                        o.channels[c].overflow == h.channels[c].overflow &&
                        o.channels[c].nonfinite == h.channels[c].nonfinite &&
                        o.channels[c].minimum == h.channels[c].minimum && o.channels[c].maximum == h.channels[c].maximum;
    }
    CHECK(partition, "histogram_rgb: independent of the tile partition");
    const double nan = std::nan(""), inf = std::numeric_limits<double>::infinity();
    bool rejected = true;
    const unsigned calls = probe.calls;
    for (const HistogramOptions& o : {HistogramOptions{0, 0, 1}, HistogramOptions{65537, 0, 1}, HistogramOptions{4, 1, 1},
                                      HistogramOptions{4, 2, 1}, HistogramOptions{4, nan, 1}, HistogramOptions{4, 0, inf},
                                      HistogramOptions{4, -1e308, 1e308}})
        rejected = rejected && throws_as([&] { histogram_rgb(probe, bounds, request, o); }, 2);
    rejected = rejected && probe.calls == calls;
    rejected = rejected && throws_as([&] { histogram_rgb(probe, bounds, RenderRequest{Rect{9, 7, 0, 1}}); }, 2);
    rejected = rejected && throws_as([&] { histogram_rgb(probe, bounds, RenderRequest{Rect{36, 26, 4, 4}}); }, 1);
    for (int f = 1; f <= 3; ++f) {
        probe.fault = f;
        rejected = rejected && throws_as([&] { histogram_rgb(probe, bounds, request); }, 2);
    }
    probe.fault = 0;
    CHECK(rejected, "histogram_rgb: invalid options (before rendering), viewports and malformed tiles rejected");
    CancellationToken pre;
    pre.cancel();
    const unsigned before = probe.calls;
    bool cancelled = throws_as([&] { histogram_rgb(probe, bounds, request, {}, &pre); }, 4) && probe.calls == before;
    CancellationToken during;
    probe.cancel_on_render = &during;
    cancelled = cancelled && throws_as([&] { histogram_rgb(probe, bounds, RenderRequest{bounds}, {}, &during); }, 4);
    probe.cancel_on_render = nullptr;
    CHECK(cancelled, "histogram_rgb: cancellation before and during the render throws");
    RawMetadata m;
    m.width = 5; m.height = 3; m.black_levels.fill(0); m.white_levels.fill(65535);
    ImageGraph g(RawImage(m, std::vector<uint16_t>(15, 32768)), GraphRecipe{});
    const auto gh = histogram_rgb(g, RenderRequest{Rect{0, 0, 5, 3}, 2});
    bool covered = gh.pixel_count == 15 && gh.descriptor == g.output().output_descriptor();
    for (const auto& c : gh.channels) {
        uint64_t n = c.underflow + c.overflow + c.nonfinite;
        for (auto k : c.counts) n += k;
        covered = covered && n == 15;
    }
    CHECK(covered, "histogram_rgb: ImageGraph overload covers every pixel on the device path");
}

int main() {
    const std::array<double, 9> identity{1, 0, 0, 0, 1, 0, 0, 0, 1};
    // 1. Bilinear linear ProPhoto through ImageGraph, 256 tiles.
    {
        RawMetadata m;
        m.width = 300; m.height = 200; m.black_levels.fill(512); m.white_levels.fill(16383);
        auto s = random_samples(300 * 200, 1, 16383);
        GraphRecipe r;
        r.red_gain = 2.0f; r.blue_gain = 1.5f;
        r.camera_color = CameraColorTransform{identity, WorkingSpace::LinearProPhotoD50};
        ImageGraph g(RawImage(m, s), r, RawDemosaicIdentity{}, true);
        std::vector<float> canvas(300 * 200 * 3);
        Renderer{}.render_tiles(g, RenderRequest{Rect{0, 0, 300, 200}, 64}, [&](const Tile& t) {
            for (uint32_t y = 0; y < t.bounds.height; ++y)
                std::memcpy(&canvas[(size_t(t.bounds.y + y) * 300 + t.bounds.x) * 3], &t.rgb[size_t(y) * t.bounds.width * 3],
                            t.bounds.width * 12);
        });
        ro_bayer b = oracle_meta(m);
        std::vector<float> exp(300 * 200 * 3);
        ro_demosaic_bilinear(&b, s.data(), {0, 0, 300, 200}, exp.data());
        const float gains[3] = {2.0f, 1.0f, 1.5f};
        ro_scale3(exp.data(), 300 * 200, gains);
        ro_scale(exp.data(), exp.size(), ro_exp2f(0.0f));
        const auto& cw = dynamic_cast<const CameraToWorkingNode&>(g.output());
        ro_mat3(exp.data(), 300 * 200, cw.matrix().data());
        CHECK(same_bits(canvas, exp), "ImageGraph bilinear linear ProPhoto, render_tiles 64 == oracle (bit-exact)");
        CHECK(same_bits(Renderer{}.render_roi(g, Rect{0, 0, 300, 200}, 1000), exp), "render_roi == oracle");
    }
    // 2. Menon node chain, odd active area, render_image of an ROI.
    {
        RawMetadata m;
        m.width = 150; m.height = 98; m.row_stride_samples = 160; m.pattern = BayerPattern::BGGR; m.cfa_phase_x = 1;
        m.active_area = {3, 2, 141, 90}; m.black_levels = {510, 512, 511, 513}; m.white_levels.fill(16383);
        auto s = random_samples(160 * 98, 2, 16383);
        RawImage img(m, s);
        auto n = std::make_shared<RawUnpackNode>(img, RawDemosaicIdentity{"rawengine.menon_base", 1});
        const Rect roi{20, 11, 90, 60};
        const Tile t = n->render(roi);
        ro_bayer b = oracle_meta(m);
        std::vector<float> exp(size_t(roi.width) * roi.height * 3);
        ro_demosaic_menon(&b, s.data(), {roi.x, roi.y, roi.width, roi.height}, exp.data());
        CHECK(same_bits(t.rgb, exp), "RawUnpackNode(Menon)::render(ROI) == oracle (bit-exact)");
        CHECK(t.descriptor == ImageDescriptor::camera_linear(), "Menon tile descriptor");
        const Rect halo = n->input_region({23, 32, 3, 4}, img.metadata().active_area);
        CHECK(halo.x == 17 && halo.y == 26 && halo.width == 15 && halo.height == 16, "Menon input_region halo 6");
    }
    // 3. LegacyBounded ImageGraph (tone + clip) and errors.
    {
        RawMetadata m;
        m.width = 64; m.height = 48; m.black_levels.fill(256); m.white_levels.fill(4095);
        auto s = random_samples(64 * 48, 3, 4095);
        GraphRecipe r;
        r.red_gain = 1.5f; r.exposure_stops = 0.5f; r.tone_shoulder = 0.3f;
        ImageGraph g(RawImage(m, s), r);
        const Tile t = Renderer{}.render_image(g, Rect{0, 0, 64, 48});
        ro_bayer b = oracle_meta(m);
        std::vector<float> exp(64 * 48 * 3);
        ro_demosaic_bilinear(&b, s.data(), {0, 0, 64, 48}, exp.data());
        const float gains[3] = {1.5f, 1.0f, 1.0f};
        ro_scale3(exp.data(), 64 * 48, gains);
        ro_scale(exp.data(), exp.size(), ro_exp2f(0.5f));
        ro_tone_curve(exp.data(), exp.size(), 0.3f, 1.0f);
        ro_clip(exp.data(), exp.size());
        CHECK(same_bits(t.rgb, exp), "ImageGraph LegacyBounded == oracle (bit-exact)");
        CHECK(t.descriptor == ImageDescriptor::bounded_output(), "bounded descriptor");
        CHECK(throws_as([&] { Renderer{}.render_image(g, Rect{60, 0, 5, 1}); }, 1), "viewport outside -> out_of_range");
        CHECK(throws_as([&] { Renderer{}.render_image(g, RenderRequest{Rect{0, 0, 4, 4}, 256, {1, RenderQuality::Preview}}); }, 2),
              "unsupported level -> invalid_argument");
        CancellationToken tok;
        tok.cancel();
        CHECK(throws_as([&] { Renderer{}.render_tiles(g, Rect{0, 0, 4, 4}, [](const Tile&) {}, 256, &tok); }, 4),
              "cancelled token -> RenderCancelled");
        CHECK(throws_as([] { RawImage(RawMetadata{}, {}); }, 2), "invalid metadata -> invalid_argument");
    }
    // 4. A foreign node feeding built-in device nodes.
    {
        auto src = std::make_shared<ConstantCamera>();
        CameraColorTransform t{identity, WorkingSpace::LinearProPhotoD50};
        CameraToWorkingNode node(src, t);
        const Tile out = node.render({0, 0, 1, 1});
        CHECK(std::abs(out.rgb[0] - 0.60399f) < 1e-3f && std::abs(out.rgb[1] - 0.1068f) < 1e-3f &&
                  std::abs(out.rgb[2] - 0.12123f) < 1e-3f, "foreign node -> CameraToWorking on device (ICC KAT)");
    }
    // 5. Guided filter + exposure + curves over a raster.
    {
        const uint32_t W = 57, H = 41;
        std::vector<float> px(W * H * 3);
        std::mt19937 rng(5);
        std::lognormal_distribution<float> ln(-2.0f, 1.5f);
        for (auto& v : px) v = std::min(ln(rng), 0.999f);
        auto raster = std::make_shared<RasterSourceNode>(RasterImage({W, H, 0, WorkingSpace::LinearRec2020D65}, px));
        auto gf = std::make_shared<WorkingYGuidedFilterNode>(raster, Rect{0, 0, W, H}, WorkingYGuidedFilterSettings{4, 1e-3});
        auto ex = std::make_shared<ExposureNode>(gf, 0.7f);
        const Rect vp{3, 2, 50, 35};
        const Tile t = Renderer{}.render_image(*ex, Rect{0, 0, W, H}, RenderRequest{vp, 16});
        const ro_rect needed{0, 0, W, H};
        std::vector<float> exp(size_t(vp.width) * vp.height * 3);
        ro_guided_filter(px.data(), needed, {vp.x, vp.y, vp.width, vp.height}, {0, 0, W, H}, 4, 1e-3,
                         0.26270021201126703, 0.059301716469861945, exp.data());
        ro_scale(exp.data(), exp.size(), ro_exp2f(0.7f));
        float worst = 0;
        for (size_t i = 0; i < exp.size(); ++i)
            worst = std::max(worst, std::abs(t.rgb[i] - exp[i]) / std::max(1.0f, std::abs(exp[i])));
        std::printf("     guided filter max rel err %.3g\n", worst);
        CHECK(worst <= 1e-5f, "GuidedFilter(r=4)+Exposure == oracle within 1e-5");
        CHECK(throws_as([&] { WorkingYGuidedFilterNode(raster, Rect{0, 0, W, H}, {9, 1e-3}); }, 2), "radius 9 rejected");
    }
    histogram_checks();
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
