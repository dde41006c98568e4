// Test-only harness over the compiled reference (tylermcm/librawops sources,
// built from /root/reference by oracle/Makefile into oracle/_ref/).  It builds
// the reference node chains for the render path and exposes them through a
// small C interface so the tests can compare the oracle restatement (and the
// HIP path) with the real reference.  Never part of the product.
#include "EditGraph.hpp"
#include "RawEngine.hpp"
#include "SpatialOps.hpp"
#include "ToneOps.hpp"

#include <atomic>
#include <chrono>
#include <exception>
#include <cstring>
#include <thread>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rawops_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace rawengine;

namespace {
thread_local std::string last_error;

int fail(const std::exception& e, int code) {
    last_error = e.what();
    return code;
}

template <class F> int guarded(F body) {
    try {
        body();
        return 0;
    } catch (const std::invalid_argument& e) { return fail(e, 1); }
    catch (const std::out_of_range& e) { return fail(e, 2); }
    catch (const std::domain_error& e) { return fail(e, 3); }
    catch (const std::length_error& e) { return fail(e, 4); }
    catch (const std::overflow_error& e) { return fail(e, 5); }
    catch (const RenderCancelled& e) { return fail(e, 6); }
    catch (const std::exception& e) { return fail(e, 7); }
}

// Emits unit-vector pixels e_0, e_1, e_2 so a matrix node's output reveals
// its float32 coefficients column by column.
class UnitColumns final : public Node {
public:
    explicit UnitColumns(ImageDescriptor d) : d_(d) {}
    Tile render(Rect b) const override {
        Tile t{b, std::vector<float>(size_t(b.width) * b.height * 3), d_};
        for (size_t i = 0; i < size_t(b.width) * b.height; ++i) t.rgb[i * 3 + (i % 3)] = 1.0f;
        return t;
    }
    ImageDescriptor output_descriptor() const noexcept override { return d_; }
private:
    ImageDescriptor d_;
};

void columns_to_rows(const std::vector<float>& cols, float out[9]) {
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) out[r * 3 + j] = cols[size_t(j) * 3 + r];
}

RawImage make_image(const ro_bayer* b, const uint16_t* samples, size_t count) {
    RawMetadata m;
    m.width = b->width; m.height = b->height; m.row_stride_samples = b->row_stride;
    m.pattern = static_cast<BayerPattern>(b->pattern);
    m.cfa_phase_x = uint8_t(b->phase_x); m.cfa_phase_y = uint8_t(b->phase_y);
    m.active_area = {b->area.x, b->area.y, b->area.width, b->area.height};
    for (int i = 0; i < 4; ++i) { m.black_levels[i] = b->black[i]; m.white_levels[i] = b->white[i]; }
    return RawImage(m, std::vector<uint16_t>(samples, samples + count));
}
} // namespace

extern "C" {

const char* refh_last_error() { return last_error.c_str(); }

// output_mode: 0 linear (ends at CameraToWorking, or Exposure without camera
// colour), 1 LegacyBounded (tone + clip), 2 SrgbPreview (to-sRGB, tone, encode).
int refh_render_raw(const ro_bayer* b, const uint16_t* samples, size_t count, int menon,
                    const float wb[3], float stops, int has_camera, const double cam[9],
                    int target, int output_mode, float shoulder, float gamma,
                    ro_rect viewport, uint32_t mip, uint32_t quality, uint32_t tile,
                    float* out) {
    return guarded([&] {
        const RawImage image = make_image(b, samples, count);
        const Rect area = image.metadata().active_area;
        RawDemosaicIdentity id;
        if (menon) id = {"rawengine.menon_base", 1};
        std::shared_ptr<const Node> n = std::make_shared<RawUnpackNode>(image, id);
        n = std::make_shared<WhiteBalanceNode>(n, wb[0], wb[1], wb[2]);
        n = std::make_shared<ExposureNode>(n, stops);
        if (has_camera) {
            CameraColorTransform t;
            for (int i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = cam[i];
            t.target = static_cast<WorkingSpace>(target);
            n = std::make_shared<CameraToWorkingNode>(n, t, area);
        }
        if (output_mode == 1) {
            n = std::make_shared<ToneCurveNode>(n, shoulder, gamma);
            n = std::make_shared<OutputClipNode>(n);
        } else if (output_mode == 2) {
            n = std::make_shared<WorkingToSrgbNode>(n);
            n = std::make_shared<ToneCurveNode>(n, shoulder, gamma);
            n = std::make_shared<SrgbEncodeNode>(n);
        }
        RenderRequest req{Rect{viewport.x, viewport.y, viewport.width, viewport.height}, tile,
                          RenderLevel{mip, static_cast<RenderQuality>(quality)}};
        const Tile t = Renderer{}.render_image(*n, area, req);
        std::memcpy(out, t.rgb.data(), t.rgb.size() * sizeof(float));
    });
}

// Wall seconds per round of `nreq` render requests served by `threads`
// worker threads, the way a TileScheduler worker pool serves them: each
// worker takes the next request and renders it (render_image of its
// viewport at its level and tile size) through the chain of refh_render_raw
// (output_mode 2: SrgbPreview) of its source, one OpenMP thread per worker;
// `reps` rounds after one untimed round.  Request k: source req_src[k],
// Menon if req_menon[k], viewport req_vp[k], mip req_mip[k] (Final at mip
// 0, Preview above), tile size req_tile[k].
int refh_time_requests(const ro_bayer* metas, const uint16_t* const* samples, const size_t* counts, uint32_t nsrc,
                       const uint32_t* req_src, const uint32_t* req_menon, const ro_rect* req_vp,
                       const uint32_t* req_mip, const uint32_t* req_tile, uint32_t nreq, const float wb[3],
                       float stops, const double cam[9], int output_mode, float shoulder, float gamma,
                       uint32_t threads, uint32_t reps, double* seconds) {
    return guarded([&] {
        std::vector<RawImage> images;
        for (uint32_t i = 0; i < nsrc; ++i) images.push_back(make_image(&metas[i], samples[i], counts[i]));
        std::vector<std::shared_ptr<const Node>> chains;   // [source][demosaic]
        for (uint32_t i = 0; i < nsrc; ++i)
            for (int menon = 0; menon < 2; ++menon) {
                RawDemosaicIdentity id;
                if (menon) id = {"rawengine.menon_base", 1};
                const Rect area = images[i].metadata().active_area;
                std::shared_ptr<const Node> n = std::make_shared<RawUnpackNode>(images[i], id);
                n = std::make_shared<WhiteBalanceNode>(n, wb[0], wb[1], wb[2]);
                n = std::make_shared<ExposureNode>(n, stops);
                CameraColorTransform t;
                for (int k = 0; k < 9; ++k) t.camera_to_xyz_d50[k] = cam[k];
                t.target = WorkingSpace::LinearProPhotoD50;
                n = std::make_shared<CameraToWorkingNode>(n, t, area);
                if (output_mode == 2) {
                    n = std::make_shared<WorkingToSrgbNode>(n);
                    n = std::make_shared<ToneCurveNode>(n, shoulder, gamma);
                    n = std::make_shared<SrgbEncodeNode>(n);
                }
                chains.push_back(n);
            }
        // A worker's exception is carried out of its thread and rethrown
        // here (the first one), after every worker has stopped.
        auto round = [&] {
            std::atomic<uint32_t> next{0};
            std::vector<std::exception_ptr> errors(threads);
            std::vector<std::thread> pool;
            for (uint32_t w = 0; w < threads; ++w)
                pool.emplace_back([&, w] {
                    try {
#ifdef _OPENMP
                        omp_set_num_threads(1);
#endif
                        for (uint32_t k = next++; k < nreq; k = next++) {
                            const uint32_t i = req_src[k];
                            if (i >= nsrc) throw std::out_of_range("request source out of range");
                            const ro_rect& v = req_vp[k];
                            const RenderRequest req{Rect{v.x, v.y, v.width, v.height}, req_tile[k],
                                                    RenderLevel{req_mip[k], req_mip[k] == 0 ? RenderQuality::Final
                                                                                            : RenderQuality::Preview}};
                            const Tile out = Renderer{}.render_image(*chains[2 * i + (req_menon[k] ? 1 : 0)],
                                                                     images[i].metadata().active_area, req);
                            if (out.rgb.empty()) throw std::runtime_error("empty render");
                        }
                    } catch (...) {
                        errors[w] = std::current_exception();
                        next = nreq;   // the other workers stop after their current request
                    }
                });
            for (auto& th : pool) th.join();
            for (auto& e : errors)
                if (e) std::rethrow_exception(e);
        };
        round();
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < reps; ++i) round();
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
    });
}

// Wall seconds per round of rendering `n` crops of one RAW source through
// Menon -> white balance -> exposure -> camera-to-ProPhoto -> working-Y
// guided filter (radius, epsilon) -> working-to-sRGB -> tone (shoulder 0.25,
// gamma 1) -> sRGB encode, the
// forward of the differentiable full chain, one crop per worker at a time
// on `threads` worker threads (one OpenMP thread each); `reps` rounds after
// one untimed round.
int refh_time_guided_crops(const ro_bayer* b, const uint16_t* samples, size_t count, const float wb[3], float stops,
                           const double cam[9], uint32_t radius, double epsilon, const ro_rect* crops, uint32_t n,
                           uint32_t threads, uint32_t reps, double* seconds) {
    return guarded([&] {
        const RawImage image = make_image(b, samples, count);
        const Rect area = image.metadata().active_area;
        std::shared_ptr<const Node> node =
            std::make_shared<RawUnpackNode>(image, RawDemosaicIdentity{"rawengine.menon_base", 1});
        node = std::make_shared<WhiteBalanceNode>(node, wb[0], wb[1], wb[2]);
        node = std::make_shared<ExposureNode>(node, stops);
        CameraColorTransform t;
        for (int k = 0; k < 9; ++k) t.camera_to_xyz_d50[k] = cam[k];
        t.target = WorkingSpace::LinearProPhotoD50;
        node = std::make_shared<CameraToWorkingNode>(node, t, area);
        node = std::make_shared<WorkingYGuidedFilterNode>(node, area, WorkingYGuidedFilterSettings{radius, epsilon});
        node = std::make_shared<WorkingToSrgbNode>(node);
        node = std::make_shared<ToneCurveNode>(node, 0.25f, 1.0f);   // display-linear, as SrgbEncodeNode requires
        node = std::make_shared<SrgbEncodeNode>(node);
        auto round = [&] {
            std::atomic<uint32_t> next{0};
            std::vector<std::exception_ptr> errors(threads);
            std::vector<std::thread> pool;
            for (uint32_t w = 0; w < threads; ++w)
                pool.emplace_back([&, w] {
                    try {
#ifdef _OPENMP
                        omp_set_num_threads(1);
#endif
                        for (uint32_t k = next++; k < n; k = next++) {
                            const ro_rect& v = crops[k];
                            const RenderRequest req{Rect{v.x, v.y, v.width, v.height}, std::max(v.width, v.height),
                                                    RenderLevel{}};
                            const Tile out = Renderer{}.render_image(*node, area, req);
                            if (out.rgb.empty()) throw std::runtime_error("empty render");
                        }
                    } catch (...) {
                        errors[w] = std::current_exception();
                        next = n;
                    }
                });
            for (auto& th : pool) th.join();
            for (auto& e : errors)
                if (e) std::rethrow_exception(e);
        };
        round();
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < reps; ++i) round();
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
    });
}

// Scene-linear raster -> guided filter -> exposure -> extended curves.
int refh_guided_chain(const float* raster, uint32_t w, uint32_t h, int space,
                      uint32_t radius, double epsilon, float stops,
                      const double* xs, const double* ys, const uint32_t offsets[5], int cubic,
                      ro_rect viewport, uint32_t mip, uint32_t quality, uint32_t tile, float* out) {
    return guarded([&] {
        RasterMetadata m{w, h, 0, static_cast<WorkingSpace>(space)};
        const RasterImage image(m, std::vector<float>(raster, raster + size_t(w) * h * 3));
        std::shared_ptr<const Node> n = std::make_shared<RasterSourceNode>(image);
        n = std::make_shared<WorkingYGuidedFilterNode>(n, Rect{0, 0, w, h},
                                                       WorkingYGuidedFilterSettings{radius, epsilon});
        if (stops != 0.0f) n = std::make_shared<ExposureNode>(n, stops);
        if (xs) {
            ExtendedCurvesSettings s;
            s.interpolation = cubic ? CurveInterpolation::ShapePreservingCubic : CurveInterpolation::Linear;
            for (int c = 0; c < 4; ++c) {
                std::vector<CurvePoint> k;
                for (uint32_t i = offsets[c]; i < offsets[c + 1]; ++i) k.push_back({xs[i], ys[i]});
                if (c == 0) s.master.knots = k; else s.channels[c - 1].knots = k;
            }
            n = std::make_shared<ExtendedCurvesNode>(n, s);
        }
        RenderRequest req{Rect{viewport.x, viewport.y, viewport.width, viewport.height}, tile,
                          RenderLevel{mip, static_cast<RenderQuality>(quality)}};
        const Tile t = Renderer{}.render_image(*n, Rect{0, 0, w, h}, req);
        std::memcpy(out, t.rgb.data(), t.rgb.size() * sizeof(float));
    });
}

// float32 coefficients the reference uses for camera->working (row-major),
// working->linear sRGB, and working->other working space.
int refh_matrices(const double cam[9], int target, float cam_to_working[9],
                  float working_to_srgb[9], float working_to_other[9]) {
    return guarded([&] {
        CameraColorTransform t;
        for (int i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = cam[i];
        t.target = static_cast<WorkingSpace>(target);
        auto cols = std::make_shared<UnitColumns>(ImageDescriptor::camera_linear());
        columns_to_rows(CameraToWorkingNode(cols, t).render({0, 0, 3, 1}).rgb, cam_to_working);
        const auto space = static_cast<WorkingSpace>(target);
        auto scene = std::make_shared<UnitColumns>(ImageDescriptor::scene_linear(space));
        columns_to_rows(WorkingToSrgbNode(scene).render({0, 0, 3, 1}).rgb, working_to_srgb);
        const auto other = space == WorkingSpace::LinearProPhotoD50 ? WorkingSpace::LinearRec2020D65
                                                                     : WorkingSpace::LinearProPhotoD50;
        columns_to_rows(WorkingSpaceConvertNode(scene, other).render({0, 0, 3, 1}).rgb, working_to_other);
    });
}

// Host-side exp2 the reference's ExposureNode uses for its multiplier.
float refh_exposure_multiplier(float stops) {
    const float one = 1.0f;
    float v = one;
    guarded([&] {
        RasterMetadata m{1, 1, 0, WorkingSpace::LinearProPhotoD50};
        auto src = std::make_shared<RasterSourceNode>(RasterImage(m, {1.0f, 1.0f, 1.0f}));
        v = ExposureNode(src, stops).render({0, 0, 1, 1}).rgb[0];
    });
    return v;
}

// Wall seconds per round of the reference's fastest legal Menon
// configuration on the host: `threads` concurrent band requests
// (render_image of one band of `rect` each, tile = band, one OpenMP thread
// per request) on one const Menon [-> WB -> exposure] -> camera-to-working
// chain built once (wb = NULL: no WB or exposure node), `reps` rounds after
// one untimed round.
int refh_time_menon_bands(const ro_bayer* b, const uint16_t* samples, size_t count, const float wb[3], float stops,
                          const double cam[9], ro_rect rect, uint32_t threads, uint32_t reps, double* seconds) {
    return guarded([&] {
        const RawImage image = make_image(b, samples, count);
        const Rect area = image.metadata().active_area;
        std::shared_ptr<const Node> n =
            std::make_shared<RawUnpackNode>(image, RawDemosaicIdentity{"rawengine.menon_base", 1});
        if (wb) {
            n = std::make_shared<WhiteBalanceNode>(n, wb[0], wb[1], wb[2]);
            n = std::make_shared<ExposureNode>(n, stops);
        }
        CameraColorTransform t;
        for (int i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = cam[i];
        t.target = WorkingSpace::LinearProPhotoD50;
        n = std::make_shared<CameraToWorkingNode>(n, t, area);
        const uint32_t band = (rect.height + threads - 1) / threads;
        auto round = [&] {
            std::vector<std::thread> pool;
            for (uint32_t k = 0; k < threads; ++k) {
                const uint32_t y0 = k * band;
                if (y0 >= rect.height) break;
                const Rect r{rect.x, rect.y + y0, rect.width, std::min(band, rect.height - y0)};
                pool.emplace_back([&, r] {
#ifdef _OPENMP
                    omp_set_num_threads(1);
#endif
                    const RenderRequest req{r, std::max(r.width, r.height), RenderLevel{}};
                    const Tile out = Renderer{}.render_image(*n, area, req);
                    if (out.rgb.empty()) throw std::runtime_error("empty band");
                });
            }
            for (auto& th : pool) th.join();
        };
        round();
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < reps; ++i) round();
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
    });
}

// ---- edit manifests (EditGraph.cpp) ----

// serialize_edit_manifest(parse_edit_manifest(json)) into out (NUL-terminated;
// *len receives the full length, also when it does not fit).
int refh_manifest_canonical(const char* json, char* out, size_t cap, size_t* len) {
    return guarded([&] {
        const std::string text = serialize_edit_manifest(parse_edit_manifest(json));
        *len = text.size();
        if (text.size() + 1 <= cap) std::memcpy(out, text.c_str(), text.size() + 1);
    });
}

int refh_fingerprint_raw(const ro_bayer* b, const uint16_t* samples, size_t count, uint8_t out[32]) {
    return guarded([&] {
        const auto d = fingerprint_raw_source(make_image(b, samples, count));
        std::memcpy(out, d.data(), 32);
    });
}

int refh_fingerprint_raster(uint32_t w, uint32_t h, uint32_t stride, int space, const float* pixels,
                            uint8_t out[32]) {
    return guarded([&] {
        const size_t n = size_t(stride ? stride : w) * h * 3;
        RasterImage image({w, h, stride, static_cast<WorkingSpace>(space)}, std::vector<float>(pixels, pixels + n));
        const auto d = fingerprint_raster_source(image);
        std::memcpy(out, d.data(), 32);
    });
}

// snapshot_legacy_recipe for a Bayer (raw != 0) or ProPhoto raster source.
int refh_legacy_snapshot(int raw, const uint8_t digest[32], const char* source_id, const char* op_id, int era,
                         const float gains[3], float stops, float shoulder, float gamma, int mode, int has_camera,
                         const double cam[9], int target, char* out, size_t cap, size_t* len) {
    return guarded([&] {
        EditSource src;
        src.id = source_id;
        src.kind = raw ? EditSourceKind::DecodedBayerU16 : EditSourceKind::SceneLinearRasterF32;
        if (!raw) src.working_space = static_cast<WorkingSpace>(target);
        std::memcpy(src.content_sha256.data(), digest, 32);
        GraphRecipe r;
        r.red_gain = gains[0]; r.green_gain = gains[1]; r.blue_gain = gains[2];
        r.exposure_stops = stops; r.tone_shoulder = shoulder; r.tone_gamma = gamma;
        r.output_mode = static_cast<OutputMode>(mode);
        if (has_camera) {
            CameraColorTransform t;
            for (int i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = cam[i];
            t.target = static_cast<WorkingSpace>(target);
            r.camera_color = t;
        }
        const std::string text = serialize_edit_manifest(
            snapshot_legacy_recipe(src, r, static_cast<LegacyRecipeEra>(era), op_id));
        *len = text.size();
        if (text.size() + 1 <= cap) std::memcpy(out, text.c_str(), text.size() + 1);
    });
}

// local_statistics_rgb on a tile covering `needed` (camera-linear descriptor).
int refh_local_statistics(const float* in, ro_rect needed, ro_rect output, ro_rect image, uint32_t radius,
                          double* mean, double* variance, uint32_t* count) {
    return guarded([&] {
        Tile t{Rect{needed.x, needed.y, needed.width, needed.height},
               std::vector<float>(in, in + size_t(needed.width) * needed.height * 3), ImageDescriptor::camera_linear()};
        const auto r = local_statistics_rgb(t, Rect{output.x, output.y, output.width, output.height},
                                            Rect{image.x, image.y, image.width, image.height}, radius);
        std::memcpy(mean, r.mean.data(), r.mean.size() * sizeof(double));
        std::memcpy(variance, r.variance.data(), r.variance.size() * sizeof(double));
        std::memcpy(count, r.finite_count.data(), r.finite_count.size() * sizeof(uint32_t));
    });
}

} // extern "C"
