// Drop-in API test: the rawengine:: C++ layer (include/rawengine/*.hpp) on the
// GPU, checked against the CPU oracle (oracle/liboracle.so) on the same
// seeded inputs.  Exit code 0 = all checks passed; prints one line per check.
#include <RawEngine.hpp>
#include <SpatialOps.hpp>
#include <ToneOps.hpp>
#include <ImageAnalysis.hpp>
#include <EditGraph.hpp>

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


// ------------------------------------------------------------ edit graphs --

static std::string uid(unsigned i) {
    char b[40];
    std::snprintf(b, sizeof b, "00000000-0000-0000-0000-%012x", i);
    return b;
}

static EditOperation make_op(unsigned id, std::string type, EditDomain in, EditDomain out, std::string upstream,
                             EditValue::Object params = {}) {
    EditOperation op;
    op.id = uid(id);
    op.type_id = std::move(type);
    op.processing_version = kCurrentEditProcessingVersion;
    op.input_domain = in;
    op.output_domain = out;
    op.inputs.emplace("image", std::move(upstream));
    op.parameters = std::move(params);
    return op;
}

static bool same_tile(const Tile& a, const Tile& b) {
    return a.descriptor == b.descriptor && a.bounds.x == b.bounds.x && a.bounds.y == b.bounds.y &&
           a.bounds.width == b.bounds.width && a.bounds.height == b.bounds.height && same_bits(a.rgb, b.rgb);
}

// Maps display-linear RGB into [0, 1] by a fixed affine map (a stand-in for an
// ICC engine: the hook only needs a deterministic host transform).
class HalfGreyDisplay final : public IccDisplayTransform {
public:
    bool unbounded = false;
    std::array<std::uint8_t, 32> profile_sha256() const noexcept override {
        std::array<std::uint8_t, 32> d{};
        d.fill(0x3a);
        return d;
    }
    std::optional<IccProfileIdentity> output_icc_identity() const override {
        IccProfileIdentity id;
        id.profile_sha256 = profile_sha256();
        return id;
    }
    void apply(float* rgb, std::size_t pixels) const override {
        for (std::size_t i = 0; i < pixels * 3; ++i)
            rgb[i] = unbounded ? 2.0f : 0.25f + 0.5f * std::min(std::max(rgb[i], 0.0f), 1.0f);
    }
};

static void edit_graph_checks() {
    const std::array<double, 9> identity{1, 0, 0, 0, 1, 0, 0, 0, 1};
    Renderer renderer;
    // (a) A RAW manifest equals the fixed ImageGraph it restates
    //     (edit_graph_tests.cpp:111-223 of the reference).
    {
        constexpr uint32_t W = 13, H = 9;
        std::vector<uint16_t> s(W * H);
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) s[y * W + x] = uint16_t(1000 + x * 1900 + y * 2300);
        RawImage raw(W, H, s);
        GraphRecipe recipe;
        recipe.red_gain = 1.15f; recipe.green_gain = 0.95f; recipe.blue_gain = 1.3f;
        recipe.exposure_stops = 0.75f; recipe.tone_shoulder = 0.35f; recipe.tone_gamma = 1.1f;
        recipe.output_mode = OutputMode::SrgbPreview;
        recipe.camera_color = CameraColorTransform{identity, WorkingSpace::LinearProPhotoD50};
        ImageGraph fixed(raw, recipe);
        EditSource source;
        source.id = uid(1);
        source.kind = EditSourceKind::DecodedBayerU16;
        source.content_sha256 = fingerprint_raw_source(raw);
        EditManifest m;
        m.sources.push_back(source);
        using D = EditDomain;
        m.operations.push_back(make_op(2, "rawengine.white_balance", D::CameraLinear, D::CameraLinear, source.id,
                                       {{"red_gain", EditValue{double(recipe.red_gain)}},
                                        {"green_gain", EditValue{double(recipe.green_gain)}},
                                        {"blue_gain", EditValue{double(recipe.blue_gain)}}}));
        m.operations.push_back(make_op(3, "rawengine.exposure", D::CameraLinear, D::CameraLinear, uid(2),
                                       {{"stops", EditValue{double(recipe.exposure_stops)}}}));
        EditValue::Array mat;
        for (double v : identity) mat.push_back(EditValue{v});
        m.operations.push_back(make_op(4, "rawengine.camera_to_working", D::CameraLinear, D::SceneLinearProPhotoD50,
                                       uid(3), {{"matrix", EditValue{mat}}}));
        m.operations.push_back(make_op(5, "rawengine.working_to_srgb", D::SceneLinearProPhotoD50, D::SceneLinearSrgb,
                                       uid(4)));
        m.operations.push_back(make_op(6, "rawengine.tone_curve", D::SceneLinearSrgb, D::DisplayLinearSrgb, uid(5),
                                       {{"shoulder", EditValue{double(recipe.tone_shoulder)}},
                                        {"gamma", EditValue{double(recipe.tone_gamma)}}}));
        m.operations.push_back(make_op(7, "rawengine.srgb_encode", D::DisplayLinearSrgb, D::DisplayEncodedSrgb, uid(6)));
        m.output_id = uid(7);
        std::reverse(m.operations.begin(), m.operations.end());
        auto node = std::make_shared<RawUnpackNode>(raw);
        ExecutableEditGraph graph(parse_edit_manifest(serialize_edit_manifest(m)), {{source, node, {0, 0, W, H}}});
        bool same = true;
        for (Rect roi : {Rect{0, 0, W, H}, Rect{2, 1, 7, 6}, Rect{12, 8, 1, 1}})
            same = same && same_tile(renderer.render_image(graph, roi, 3), renderer.render_image(fixed, roi, 3));
        CHECK(same, "EditGraph: RAW manifest render == fixed ImageGraph (bit-exact, 3 ROIs)");
        std::vector<Tile> tiles;
        renderer.render_tiles(graph, RenderRequest{Rect{0, 0, W, H}, 4}, [&](const Tile& t) { tiles.push_back(t); });
        CHECK(tiles.size() == 12, "EditGraph: render_tiles keeps the tile partition");
        bool rejected = true;
        auto wrong = source;
        wrong.content_sha256[0] ^= 1;
        rejected &= throws_as([&] { ExecutableEditGraph(m, {{wrong, node, {0, 0, W, H}}}); }, 2);
        auto altered = raw.samples();
        altered[0] ^= 1;
        rejected &= throws_as([&] {
            ExecutableEditGraph(m, {{source, std::make_shared<RawUnpackNode>(RawImage(raw.metadata(), altered)),
                                     {0, 0, W, H}}});
        }, 2);
        rejected &= throws_as([&] { ExecutableEditGraph(m, {{source, node, {0, 0, W - 1, H}}}); }, 2);
        auto wrong_domain = m;
        wrong_domain.operations.front().input_domain = D::CameraLinear;
        rejected &= throws_as([&] { ExecutableEditGraph(wrong_domain, {{source, node, {0, 0, W, H}}}); }, 2);
        auto unknown = m;
        unknown.operations.front().type_id = "vendor.unknown.operation";
        rejected &= serialize_edit_manifest(parse_edit_manifest(serialize_edit_manifest(unknown))) ==
                    serialize_edit_manifest(unknown);
        rejected &= throws_as([&] { ExecutableEditGraph(unknown, {{source, node, {0, 0, W, H}}}); }, 2);
        auto version = m;
        version.operations.front().processing_version = kLegacyRec2020ProcessingVersion;
        rejected &= throws_as([&] { ExecutableEditGraph(version, {{source, node, {0, 0, W, H}}}); }, 2);
        CHECK(rejected, "EditGraph: wrong fingerprint/samples/bounds/domain, unknown op, mixed versions rejected");
        auto old_recipe = recipe;
        old_recipe.camera_color->target = WorkingSpace::LinearRec2020D65;
        ImageGraph old_fixed(raw, old_recipe);
        auto old_manifest = snapshot_legacy_recipe(source, old_recipe, LegacyRecipeEra::ImplicitRec2020, uid(8));
        ExecutableEditGraph old_graph(parse_edit_manifest(serialize_edit_manifest(old_manifest)),
                                      {{source, node, {0, 0, W, H}}});
        CHECK(same_tile(renderer.render_image(old_graph, Rect{1, 2, 8, 5}, 3),
                        renderer.render_image(old_fixed, Rect{1, 2, 8, 5}, 3)),
              "EditGraph: legacy Rec.2020 snapshot replay == fixed graph");
        const Rect region = old_graph.required_source_region({1, 2, 8, 5});
        CHECK(region.x == 0 && region.y == 1 && region.width == 10 && region.height == 7,
              "EditGraph: required_source_region plans the demosaic halo (edit_graph_tests.cpp:219-222)");
    }
    // (b) RAW (Menon) -> WB -> camera -> guided filter -> exposure -> curves:
    //     two device launches, checked against the oracle.
    {
        RawMetadata md;
        md.width = 300; md.height = 200; md.row_stride_samples = 304; md.pattern = BayerPattern::BGGR;
        md.cfa_phase_x = 1; md.active_area = {5, 3, 288, 190};
        md.black_levels = {510, 512, 511, 513}; md.white_levels.fill(16383);
        auto s = random_samples(304 * 200, 7, 16383);
        RawImage raw(md, s);
        EditSource source;
        source.id = uid(0x40);
        source.kind = EditSourceKind::DecodedBayerU16;
        source.content_sha256 = fingerprint_raw_source(raw);
        source.demosaic = RawDemosaicIdentity{"rawengine.menon_base", 1};
        const std::array<double, 9> cam{0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88};
        EditValue::Array mat;
        for (double v : cam) mat.push_back(EditValue{v});
        using D = EditDomain;
        auto knots = [](std::initializer_list<std::pair<double, double>> k) {
            EditValue::Array a;
            for (auto [x, y] : k) a.push_back(EditValue{EditValue::Array{EditValue{x}, EditValue{y}}});
            return EditValue{a};
        };
        EditManifest m;
        m.format_version = 3;
        m.sources.push_back(source);
        m.operations.push_back(make_op(0x41, "rawengine.white_balance", D::CameraLinear, D::CameraLinear, source.id,
                                       {{"red_gain", EditValue{2.0}}, {"green_gain", EditValue{1.0}},
                                        {"blue_gain", EditValue{1.5}}}));
        m.operations.push_back(make_op(0x42, "rawengine.camera_to_working", D::CameraLinear,
                                       D::SceneLinearProPhotoD50, uid(0x41), {{"matrix", EditValue{mat}}}));
        m.operations.push_back(make_op(0x43, "rawengine.guided_filter_working_y", D::SceneLinearProPhotoD50,
                                       D::SceneLinearProPhotoD50, uid(0x42),
                                       {{"radius", EditValue{std::int64_t(4)}}, {"epsilon", EditValue{1e-3}}}));
        m.operations.push_back(make_op(0x44, "rawengine.exposure", D::SceneLinearProPhotoD50, D::SceneLinearProPhotoD50,
                                       uid(0x43), {{"stops", EditValue{0.7}}}));
        m.operations.push_back(make_op(0x45, "rawengine.curves_extended", D::SceneLinearProPhotoD50,
                                       D::SceneLinearProPhotoD50, uid(0x44),
                                       {{"master", knots({{0, 0}, {0.25, 0.3}, {1, 1}})},
                                        {"red", knots({{0, 0}, {1, 1}})}, {"green", knots({{0, 0}, {1, 1}})},
                                        {"blue", knots({{0, 0.1}, {0.5, 0.45}, {1, 1}})},
                                        {"interpolation", EditValue{std::string("shape_preserving_cubic")}}}));
        m.output_id = uid(0x45);
        auto node = std::make_shared<RawUnpackNode>(raw, *source.demosaic);
        ExecutableEditGraph graph(m, {{source, node, md.active_area}});
        const Rect vp{40, 30, 200, 120};
        const Tile got = renderer.render_image(graph, RenderRequest{vp, 64});
        // Oracle: the guided filter's input region rendered by the restatement.
        const Rect area = md.active_area;
        const Rect needed = working_y_guided_filter_region(vp, area, 4);
        ro_bayer b = oracle_meta(md);
        std::vector<float> lin(size_t(needed.width) * needed.height * 3);
        ro_demosaic_menon(&b, s.data(), {needed.x, needed.y, needed.width, needed.height}, lin.data());
        const float gains[3] = {2.0f, 1.0f, 1.5f};
        ro_scale3(lin.data(), lin.size() / 3, gains);
        CameraToWorkingNode cw(std::make_shared<ConstantCamera>(), CameraColorTransform{cam, WorkingSpace::LinearProPhotoD50});
        ro_mat3(lin.data(), lin.size() / 3, cw.matrix().data());
        std::vector<float> exp(size_t(vp.width) * vp.height * 3);
        ro_guided_filter(lin.data(), {needed.x, needed.y, needed.width, needed.height},
                         {vp.x, vp.y, vp.width, vp.height}, {area.x, area.y, area.width, area.height}, 4, 1e-3,
                         0.28807112822929337, 0.00008565396060525903, exp.data());
        ro_scale(exp.data(), exp.size(), ro_exp2f(0.7f));
        const double xs[] = {0, 0.25, 1, 0, 1, 0, 1, 0, 0.5, 1}, ys[] = {0, 0.3, 1, 0, 1, 0, 1, 0.1, 0.45, 1};
        const uint32_t offs[5] = {0, 3, 5, 7, 10};
        ro_curves(exp.data(), exp.size() / 3, xs, ys, offs, 1);
        float worst = 0;
        for (size_t i = 0; i < exp.size(); ++i)
            worst = std::max(worst, std::abs(got.rgb[i] - exp[i]) / std::max(1.0f, std::abs(exp[i])));
        std::printf("     manifest Menon+GF+curves max rel err %.3g\n", worst);
        CHECK(got.rgb.size() == exp.size() && worst <= 1e-5f,
              "EditGraph: RAW Menon -> WB -> camera -> guided filter -> exposure -> curves == oracle within 1e-5");
        // Tile cache: warm renders hit, edits never reuse stale tiles.
        auto cache = std::make_shared<TileCache>(64u << 20);
        ExecutableEditGraph cached(m, {{source, node, md.active_area}}, nullptr, cache);
        const Tile cold = renderer.render_image(cached, RenderRequest{vp, 64});
        const auto c0 = cache->stats();
        const Tile warm = renderer.render_image(cached, RenderRequest{vp, 64});
        const auto c1 = cache->stats();
        CHECK(same_tile(cold, got) && same_tile(warm, got) && c1.hits > c0.hits && c1.misses == c0.misses &&
                  c0.entries >= 2 && c1.used_bytes <= (64u << 20),
              "TileCache: cold and warm cached renders == uncached; warm render hits");
        auto late = m;
        late.operations.back().parameters["interpolation"] = EditValue{std::string("linear")};
        ExecutableEditGraph late_graph(late, {{source, node, md.active_area}}, nullptr, cache);
        const Tile late_out = renderer.render_image(late_graph, RenderRequest{vp, 64});
        const auto c2 = cache->stats();
        CHECK(same_tile(late_out, renderer.render_image(ExecutableEditGraph(late, {{source, node, md.active_area}}),
                                                        RenderRequest{vp, 64})) &&
                  !same_bits(late_out.rgb, got.rgb) && c2.hits > c1.hits,
              "TileCache: a late edit reuses the unchanged guided-filter input tile and no stale output");
        auto early = m;
        early.operations.front().parameters["red_gain"] = EditValue{2.5};
        ExecutableEditGraph early_graph(early, {{source, node, md.active_area}}, nullptr, cache);
        CHECK(same_tile(renderer.render_image(early_graph, RenderRequest{vp, 64}),
                        renderer.render_image(ExecutableEditGraph(early, {{source, node, md.active_area}}),
                                              RenderRequest{vp, 64})),
              "TileCache: an early edit reuses no stale downstream tile");
        auto tiny = std::make_shared<TileCache>(700);
        ExecutableEditGraph tiny_graph(m, {{source, node, md.active_area}}, nullptr, tiny);
        CHECK(same_tile(renderer.render_image(tiny_graph, RenderRequest{vp, 64}), got) && tiny->stats().used_bytes <= 700,
              "TileCache: a tile larger than the budget renders without entering the cache");
        CancellationToken pre;
        pre.cancel();
        CHECK(throws_as([&] { renderer.render_image(cached, RenderRequest{vp, 64}, &pre); }, 4),
              "TileCache: a cancelled cached render throws RenderCancelled");
        cache->clear();
        CHECK(cache->stats().entries == 0 && cache->stats().used_bytes == 0 && cache->stats().hits == 0,
              "TileCache: clear() drops entries and counters");
        // Host-tile API on a source node (EditGraph tests: level keys).
        TileCache level_cache(4096);
        std::array<std::uint8_t, 32> sig{};
        sig[0] = 17;
        const Rect one{0, 0, 1, 1};
        auto raster = std::make_shared<RasterSourceNode>(RasterImage({8, 6, 0, WorkingSpace::LinearProPhotoD50},
                                                                     std::vector<float>(8 * 6 * 3, 0.25f)));
        const Tile first = level_cache.render(*raster, sig, one);
        const Tile again = level_cache.render(*raster, sig, one);
        bool levels = level_cache.stats().entries == 1 && level_cache.stats().hits == 1 && same_tile(first, again) &&
                      first.rgb == std::vector<float>{0.25f, 0.25f, 0.25f};
        level_cache.render(*raster, sig, one, {0, RenderQuality::Preview});
        level_cache.render(*raster, sig, one, {1, RenderQuality::Preview});
        levels = levels && level_cache.stats().entries == 3 && level_cache.stats().misses == 3;
        levels = levels && throws_as([&] { level_cache.render(*raster, sig, one, {32, RenderQuality::Final}); }, 2);
        CHECK(levels, "TileCache::render: one entry per (signature, bounds, mip, quality); invalid level rejected");
        // histogram_rgb over the manifest graph's output bounds.
        const auto h = histogram_rgb(graph, RenderRequest{vp, 64}, HistogramOptions{16, 0.0, 1.0});
        uint64_t total = 0;
        for (auto k : h.channels[1].counts) total += k;
        total += h.channels[1].underflow + h.channels[1].overflow + h.channels[1].nonfinite;
        CHECK(total == uint64_t(vp.width) * vp.height && h.descriptor == graph.output().output_descriptor(),
              "histogram_rgb(ExecutableEditGraph) covers the viewport");
    }
    // (c) ICC display hook: device render up to display-linear sRGB, host transform.
    {
        RawMetadata md;
        md.width = 40; md.height = 30; md.black_levels.fill(64); md.white_levels.fill(4095);
        RawImage raw(md, random_samples(40 * 30, 9, 4095));
        GraphRecipe r;
        r.camera_color = CameraColorTransform{identity, WorkingSpace::LinearProPhotoD50};
        r.output_mode = OutputMode::IccDisplay;
        auto display = std::make_shared<HalfGreyDisplay>();
        ImageGraph icc_graph(raw, r, display);
        GraphRecipe lin = r;
        lin.output_mode = OutputMode::SrgbPreview;
        ImageGraph preview(raw, lin);
        const auto& encode = dynamic_cast<const SrgbEncodeNode&>(preview.output());
        Tile tone = encode.input_node()->render({0, 0, 40, 30});
        display->apply(tone.rgb.data(), tone.rgb.size() / 3);
        const Tile got = renderer.render_image(icc_graph, Rect{0, 0, 40, 30});
        CHECK(same_bits(got.rgb, tone.rgb) && got.descriptor == ImageDescriptor::icc_display(display->profile_sha256()),
              "ICC hook: IccDisplayNode == display-linear device render + host transform");
        display->unbounded = true;
        CHECK(throws_as([&] { renderer.render_image(icc_graph, Rect{0, 0, 4, 4}); }, 3),
              "ICC hook: unbounded transform output -> domain_error");
        display->unbounded = false;
        EditSource source;
        source.id = uid(0x60);
        source.kind = EditSourceKind::DecodedBayerU16;
        source.content_sha256 = fingerprint_raw_source(raw);
        auto snapshot = snapshot_legacy_recipe(source, r, LegacyRecipeEra::ImplicitProPhoto, uid(0x61),
                                               *display->output_icc_identity());
        ExecutableEditGraph g(snapshot, {{source, std::make_shared<RawUnpackNode>(raw), {0, 0, 40, 30}}}, display);
        CHECK(same_tile(renderer.render_image(g, Rect{0, 0, 40, 30}), got),
              "ICC hook: legacy icc_display manifest == ImageGraph IccDisplay");
        CHECK(throws_as([&] { ExecutableEditGraph(snapshot, {{source, std::make_shared<RawUnpackNode>(raw),
                                                              {0, 0, 40, 30}}}); }, 2),
              "ICC hook: manifest with an output profile needs the runtime transform");
    }
}

// ------------------------------------------------------- local statistics --

// R = x - 12, G = y - 10, B = 2 (or NaN): the reference test's probe source.
class AffineSource final : public Node {
public:
    bool malformed = false, nonfinite = false;
    mutable Rect last{};
    ImageDescriptor output_descriptor() const noexcept override {
        return ImageDescriptor::scene_linear(WorkingSpace::LinearProPhotoD50);
    }
    Tile render(Rect r) const override {
        last = r;
        Tile t{r, {}, output_descriptor()};
        for (uint32_t y = 0; y < r.height; ++y)
            for (uint32_t x = 0; x < r.width; ++x) {
                t.rgb.push_back(float(int64_t(r.x) + x - 12));
                t.rgb.push_back(float(int64_t(r.y) + y - 10));
                t.rgb.push_back(nonfinite ? std::nanf("") : 2.0f);
            }
        if (malformed) t.rgb.pop_back();
        return t;
    }
};

static LocalStatisticsTile collect(const Node& node, Rect image, uint32_t size, unsigned radius) {
    const size_t n = size_t(image.width) * image.height * 3;
    LocalStatisticsTile out{image, node.output_descriptor(), radius, std::vector<double>(n), std::vector<double>(n),
                            std::vector<uint32_t>(n)};
    analyze_local_rgb(node, image, RenderRequest{image, size}, [&](const LocalStatisticsTile& t) {
        for (uint32_t y = 0; y < t.bounds.height; ++y)
            for (uint32_t x = 0; x < t.bounds.width; ++x)
                for (int c = 0; c < 3; ++c) {
                    const size_t src = (size_t(y) * t.bounds.width + x) * 3 + c;
                    const size_t dst = ((size_t(t.bounds.y - image.y) + y) * image.width + t.bounds.x - image.x + x) * 3 + c;
                    out.mean[dst] = t.mean[src];
                    out.variance[dst] = t.variance[src];
                    out.finite_count[dst] = t.finite_count[src];
                }
    }, radius);
    return out;
}

static bool same_doubles(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && !std::memcmp(a.data(), b.data(), a.size() * sizeof(double));
}

static void local_statistics_checks() {
    AffineSource source;
    const Rect image{10, 7, 7, 5};
    bool truth = true, tiles = true;
    for (unsigned radius : {0u, 1u, 3u, 8u}) {
        const auto full = collect(source, image, 256, radius);
        for (unsigned y = 0; y < 5; ++y)
            for (unsigned x = 0; x < 7; ++x) {
                const unsigned left = x > radius ? x - radius : 0, right = std::min(6u, x + radius);
                const unsigned top = y > radius ? y - radius : 0, bottom = std::min(4u, y + radius);
                const size_t i = (size_t(y) * 7 + x) * 3;
                const double nx = right - left + 1, ny = bottom - top + 1;
                truth = truth && std::abs(full.mean[i] - (double(left + right) / 2 - 2)) < 1e-12 &&
                        std::abs(full.mean[i + 1] - (double(top + bottom) / 2 - 3)) < 1e-12 &&
                        std::abs(full.variance[i] - (nx * nx - 1) / 12) < 1e-12 &&
                        std::abs(full.variance[i + 1] - (ny * ny - 1) / 12) < 1e-12 && full.mean[i + 2] == 2 &&
                        full.variance[i + 2] == 0 && full.finite_count[i] == nx * ny;
            }
        for (unsigned size : {1u, 2u, 4u}) {
            const auto tiled = collect(source, image, size, radius);
            tiles = tiles && same_doubles(tiled.mean, full.mean) && same_doubles(tiled.variance, full.variance) &&
                    tiled.finite_count == full.finite_count;
        }
    }
    CHECK(truth, "local statistics: affine means/variances and window counts (spatial_ops_tests.cpp:49-62)");
    CHECK(tiles, "local statistics: tile sizes 1/2/4 equal one 256 tile bit for bit");
    bool halo = false;
    analyze_local_rgb(source, image, RenderRequest{{12, 9, 1, 1}, 1}, [&](const LocalStatisticsTile& t) {
        halo = t.finite_count[0] == 9 && t.mean[0] == 0 && std::abs(t.variance[0] - 2.0 / 3) < 1e-12;
    }, 1);
    CHECK(halo && source.last.width == 3 && source.last.height == 3, "local statistics: interior ROI reads its halo");
    source.nonfinite = true;
    const auto missing = collect(source, image, 3, 1);
    bool empty = true;
    for (size_t i = 2; i < missing.mean.size(); i += 3)
        empty = empty && missing.finite_count[i] == 0 && std::isnan(missing.mean[i]) && std::isnan(missing.variance[i]);
    source.nonfinite = false;
    CHECK(empty, "local statistics: windows without finite samples have NaN moments");
    bool rejected = throws_as([&] { local_statistics_region(image, image, 9); }, 2) &&
                    throws_as([&] { analyze_local_rgb(source, image, RenderRequest{image, 0}, [](const auto&) {}); }, 2) &&
                    throws_as([&] { analyze_local_rgb(source, image, RenderRequest{image}, {}); }, 2) &&
                    throws_as([&] { local_statistics_region({9, 7, 1, 1}, image, 1); }, 1);
    source.malformed = true;
    rejected = rejected && throws_as([&] { collect(source, image, 3, 1); }, 2);
    source.malformed = false;
    CancellationToken pre;
    pre.cancel();
    rejected = rejected && throws_as([&] { analyze_local_rgb(source, image, RenderRequest{image}, [](const auto&) {}, 1, &pre); }, 4);
    CancellationToken last;
    unsigned callbacks = 0;
    rejected = rejected && throws_as([&] {
        analyze_local_rgb(source, image, RenderRequest{image}, [&](const auto&) { ++callbacks; last.cancel(); }, 1, &last);
    }, 4) && callbacks == 1;
    CHECK(rejected, "local statistics: validation and cancellation (spatial_ops_tests.cpp:124-138)");
    // A device chain: Menon render with its halo, moments on the device, equal to the oracle.
    RawMetadata m;
    m.width = 160; m.height = 120; m.pattern = BayerPattern::GBRG; m.black_levels.fill(256); m.white_levels.fill(4095);
    auto s = random_samples(160 * 120, 13, 4095);
    ImageGraph g(RawImage(m, s), GraphRecipe{}, RawDemosaicIdentity{"rawengine.menon_base", 1}, true);
    const Rect img{0, 0, 160, 120};
    const auto got = collect(g.output(), img, 48, 3);
    ro_bayer b = oracle_meta(m);
    std::vector<float> lin(160 * 120 * 3);
    ro_demosaic_menon(&b, s.data(), {0, 0, 160, 120}, lin.data());
    std::vector<double> mean(lin.size()), var(lin.size());
    std::vector<uint32_t> cnt(lin.size());
    ro_local_statistics(lin.data(), {0, 0, 160, 120}, {0, 0, 160, 120}, {0, 0, 160, 120}, 3, mean.data(), var.data(),
                        cnt.data());
    CHECK(same_doubles(got.mean, mean) && same_doubles(got.variance, var) && got.finite_count == cnt,
          "analyze_local_rgb(ImageGraph Menon, 48-pixel tiles, r=3) == oracle bit for bit");
}

int main() {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);  // a line per check, also when a later check aborts
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
    edit_graph_checks();
    local_statistics_checks();
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
