// C ABI of the MI355X render path (declared in include/rawops_hip.h).
// Validation mirrors the reference's constructors and renderer checks so that
// the C++ drop-in layer can rethrow the same exception types.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rawops_hip.h"
#include "color.hpp"
#include "render_launch.hpp"
#include "batch_launch.hpp"

#include <algorithm>
#include <numeric>

using namespace rawops;

struct rawops_source {
    bool is_raw = true;
    rawops_raw_metadata meta{};     // RAW: normalised metadata
    const uint16_t* samples = nullptr;
    uint32_t width = 0, height = 0, stride = 0, space = 0;  // raster
    const float* pixels = nullptr;
    void* owned = nullptr;          // device allocation owned by the source
};

struct rawops_plan {
    uint32_t demosaic = 0, mip_anchor = 0;
    OpList native{}, reduced{};
    OpList* device = nullptr;       // device copies: [0] native, [1] reduced, [2] empty
    uint32_t curve_cubic = 0;       // interpolation of the first curve op
    std::vector<void*> tables;      // device curve tables
    ~rawops_plan() {
        for (void* t : tables) (void)hipFree(t);
        if (device) (void)hipFree(device);
    }
};

struct rawops_batch {
    std::vector<rawops::BatchItem> items;      // served order
    std::vector<size_t> order;                 // served position -> request index
    rawops::BatchItem* dev_items = nullptr;
    void* state = nullptr;
    uint32_t* remaining = nullptr;
    unsigned long long* done = nullptr;
    uint32_t* flags = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int shape = 0, total = 0;
    double tuned_ms = 0.0;
    ~rawops_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {static_cast<void*>(dev_items), state, static_cast<void*>(remaining), static_cast<void*>(done),
                        static_cast<void*>(flags)})
            if (p) (void)hipFree(p);
    }
};

namespace {

thread_local std::string g_error;

struct Failure {
    rawops_status status;
    std::string message;
};

[[noreturn]] void fail(rawops_status s, const std::string& m) { throw Failure{s, m}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(RAWOPS_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
rawops_status guarded(F&& body) {
    try {
        body();
        g_error.clear();
        return RAWOPS_OK;
    } catch (const Failure& f) {
        g_error = f.message;
        return f.status;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return RAWOPS_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return RAWOPS_LENGTH_ERROR;
    } catch (const std::exception& e) {
        g_error = e.what();
        return RAWOPS_LOGIC_ERROR;
    }
}

// RawImage ctor checks, RawEngine.cpp:223-250.
rawops_raw_metadata normalise(const rawops_raw_metadata* in, size_t count) {
    if (!in) fail(RAWOPS_INVALID_ARGUMENT, "RAW metadata is required");
    rawops_raw_metadata m = *in;
    if (!m.width || !m.height || m.pattern > 3 || m.cfa_phase_x > 1 || m.cfa_phase_y > 1)
        fail(RAWOPS_INVALID_ARGUMENT, "invalid RAW dimensions, Bayer pattern, or CFA phase");
    if (m.row_stride_samples == 0) m.row_stride_samples = m.width;
    if (m.row_stride_samples < m.width) fail(RAWOPS_INVALID_ARGUMENT, "RAW row stride is smaller than width");
    const rawops_rect a0 = m.active_area;
    if (!a0.x && !a0.y && !a0.width && !a0.height) m.active_area = {0, 0, m.width, m.height};
    const rawops_rect a = m.active_area;
    if (!a.width || !a.height || uint64_t(a.x) + a.width > m.width || uint64_t(a.y) + a.height > m.height)
        fail(RAWOPS_INVALID_ARGUMENT, "RAW active area is outside the sensor");
    for (int s = 0; s < 4; ++s)
        if (m.white_levels[s] <= m.black_levels[s])
            fail(RAWOPS_INVALID_ARGUMENT, "RAW white level must exceed black level at every site");
    if (count != size_t(m.row_stride_samples) * m.height)
        fail(RAWOPS_INVALID_ARGUMENT, "Bayer sample count does not match stride and height");
    return m;
}

// RN(1/d) in float: the float nearest to 1/d (the product y * d is exact in
// double for d < 2^16, so the comparison is exact).
float nearest_reciprocal(uint32_t d) {
    const float y = float(1.0 / double(d));
    float best = y;
    double err = std::abs(double(y) * d - 1.0);
    for (const float c : {std::nextafter(y, 0.0f), std::nextafter(y, 2.0f)}) {
        const double e = std::abs(double(c) * d - 1.0);
        if (e < err) { err = e; best = c; }
    }
    return best;
}

SensorArgs sensor_args(const rawops_source& src) {
    static const int cells[4][4] = {{0, 1, 1, 2}, {2, 1, 1, 0}, {1, 0, 2, 1}, {1, 2, 0, 1}};
    const auto& m = src.meta;
    SensorArgs s{};
    s.samples = src.samples;
    s.stride = m.row_stride_samples;
    s.ax = int(m.active_area.x); s.ay = int(m.active_area.y);
    s.aw = int(m.active_area.width); s.ah = int(m.active_area.height);
    s.phase_x = m.cfa_phase_x; s.phase_y = m.cfa_phase_y;
    for (uint32_t site = 0; site < 4; ++site) s.colors |= uint32_t(cells[m.pattern][site]) << (2 * site);
    for (uint32_t b = 0; b < 2; ++b) {
        const uint32_t row_site = ((b + m.cfa_phase_y) & 1u) * 2u;
        const uint32_t site0 = row_site + ((0u + m.cfa_phase_x) & 1u);
        const bool even_is_green = cells[m.pattern][site0] == 1;
        s.ng_parity |= (even_is_green ? 1u : 0u) << b;
    }
    for (int site = 0; site < 4; ++site) {
        s.black[site] = float(m.black_levels[site]);
        s.denom[site] = float(int(m.white_levels[site]) - int(m.black_levels[site]));
        s.rcp[site] = nearest_reciprocal(uint32_t(int(m.white_levels[site]) - int(m.black_levels[site])));
    }
    return s;
}

// validate_curve, ToneOps.cpp:17-32, and the ExtendedCurvesNode ctor tables
// (ToneOps.cpp:795-816).
void prepare_curve(DevCurve& c, const double* xs, const double* ys, uint32_t n, bool cubic) {
    if (n < 2 || n > uint32_t(kMaxKnots))
        fail(RAWOPS_INVALID_ARGUMENT, "curve knot count exceeds selected operation bounds");
    for (uint32_t i = 0; i < n; ++i) {
        if (!std::isfinite(xs[i]) || !std::isfinite(ys[i]) || std::abs(xs[i]) > 65536 || std::abs(ys[i]) > 65536)
            fail(RAWOPS_INVALID_ARGUMENT, "curve coordinates must be finite with magnitude <=65536");
        if (i) {
            if (!(xs[i] > xs[i - 1])) fail(RAWOPS_INVALID_ARGUMENT, "curve x coordinates must increase strictly");
            const double slope = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
            if (!std::isfinite(slope) || std::abs(slope) > 65536)
                fail(RAWOPS_INVALID_ARGUMENT, "curve slope must be finite with magnitude <=65536");
        }
    }
    std::memset(&c, 0, sizeof(c));
    c.n = n;
    c.identity = 1;
    for (uint32_t i = 0; i < n; ++i) {
        c.x[i] = xs[i];
        c.y[i] = ys[i];
        if (xs[i] != ys[i]) c.identity = 0;
    }
    for (uint32_t i = 1; i < n; ++i) c.slope[i - 1] = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
    if (!cubic) return;
    std::vector<double> tangent(n);
    tangent[0] = c.slope[0];
    for (uint32_t i = 1; i + 1 < n; ++i) {
        const double a = c.slope[i - 1], b = c.slope[i];
        tangent[i] = a == 0 || b == 0 || std::signbit(a) != std::signbit(b)
                         ? 0.0 : std::copysign(std::min(std::abs(a), std::abs(b)), a);
    }
    tangent[n - 1] = c.slope[n - 2];
    for (uint32_t i = 0; i + 1 < n; ++i) {
        const double h = xs[i + 1] - xs[i];
        c.c0[i] = ys[i] + (h * tangent[i]) / 3;
        c.c1[i] = ys[i + 1] - (h * tangent[i + 1]) / 3;
    }
}

void compile_ops(rawops_plan& plan, const rawops_op_desc* ops, uint32_t n, OpList& out) {
    if (n > uint32_t(kMaxOps)) fail(RAWOPS_INVALID_ARGUMENT, "op program is too long");
    if (n && !ops) fail(RAWOPS_INVALID_ARGUMENT, "op program is null");
    out.n = n;
    for (uint32_t i = 0; i < n; ++i) {
        const rawops_op_desc& d = ops[i];
        DevOp& o = out.op[i];
        o.kind = d.kind;
        std::memcpy(o.v, d.v, sizeof(o.v));
        o.curves = nullptr;
        switch (d.kind) {
        case OP_GAINS:
            for (int c = 0; c < 3; ++c)
                if (!std::isfinite(d.v[c]) || d.v[c] <= 0.0f || d.v[c] > 65536.0f)
                    fail(RAWOPS_INVALID_ARGUMENT, "white balance gains must be finite and in (0, 65536]");
            break;
        case OP_SCALE:
        case OP_MATRIX:
        case OP_CLIP:
        case OP_SRGB_ENCODE:
            break;
        case OP_TONE:
            if (!std::isfinite(d.v[0]) || d.v[0] <= 0.0f || !std::isfinite(d.v[1]) || d.v[1] <= 0.0f)
                fail(RAWOPS_INVALID_ARGUMENT, "tone shoulder and gamma must be finite and positive");
            break;
        case OP_CURVES: {
            if (!d.knots_x || !d.knots_y) fail(RAWOPS_INVALID_ARGUMENT, "curve knots are required");
            auto host = std::make_unique<DevCurves>();
            host->cubic = d.cubic ? 1u : 0u;
            for (int c = 0; c < 4; ++c) {
                const uint32_t b = d.knot_offsets[c], e = d.knot_offsets[c + 1];
                if (e < b) fail(RAWOPS_INVALID_ARGUMENT, "curve knot offsets must be nondecreasing");
                prepare_curve(host->c[c], d.knots_x + b, d.knots_y + b, e - b, d.cubic != 0);
            }
            void* dev = nullptr;
            hip_check(hipMalloc(&dev, sizeof(DevCurves)), "hipMalloc(curves)");
            plan.tables.push_back(dev);
            hip_check(hipMemcpy(dev, host.get(), sizeof(DevCurves), hipMemcpyHostToDevice), "hipMemcpy(curves)");
            o.curves = static_cast<const DevCurves*>(dev);
            o.knots = host->c[0].n;
            plan.curve_cubic = d.cubic ? 1u : 0u;
            break;
        }
        default:
            fail(RAWOPS_INVALID_ARGUMENT, "unknown op kind");
        }
    }
}

rawops_rect level_bounds(rawops_rect source, uint32_t mip, uint32_t quality) {
    if (mip == 0 && (quality == RAWOPS_QUALITY_FINAL || quality == RAWOPS_QUALITY_PREVIEW)) return source;
    if (mip >= 1 && mip <= 2 && quality == RAWOPS_QUALITY_PREVIEW) {
        const uint32_t s = 1u << mip;
        return {0, 0, source.width / s + (source.width % s != 0), source.height / s + (source.height % s != 0)};
    }
    fail(RAWOPS_INVALID_ARGUMENT, "render level is not implemented for these source bounds");
}

void inside_or_throw(rawops_rect area, rawops_rect r) {
    if (r.x < area.x || r.y < area.y || uint64_t(r.x) + r.width > uint64_t(area.x) + area.width ||
        uint64_t(r.y) + r.height > uint64_t(area.y) + area.height)
        fail(RAWOPS_OUT_OF_RANGE, "viewport is outside the active RAW area");
}

void validate_render(const rawops_plan* plan, const rawops_source* src, rawops_rect viewport, uint32_t mip,
                     uint32_t quality) {
    if (!plan || !src) fail(RAWOPS_INVALID_ARGUMENT, "plan and source are required");
    if (!src->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "render needs a RAW source");
    inside_or_throw(level_bounds(src->meta.active_area, mip, quality), viewport);
    if (mip == 0 && quality != RAWOPS_QUALITY_FINAL)
        fail(RAWOPS_INVALID_ARGUMENT, "node does not support this render level");
    if (mip > 0 && !plan->mip_anchor)
        fail(RAWOPS_INVALID_ARGUMENT, "node does not support this render level");
}

}  // namespace

extern "C" {

const char* rawops_last_error(void) { return g_error.c_str(); }
const char* rawops_version(void) { return "rawops-mi355x 0.1 (gfx950)"; }

rawops_status rawops_camera_to_working_matrix(const double cam[9], uint32_t space, float out[9]) {
    return guarded([&] {
        if (!cam || !out) fail(RAWOPS_INVALID_ARGUMENT, "null matrix pointer");
        Mat3 m;
        for (int i = 0; i < 9; ++i) m[i] = cam[i];
        const auto f = camera_to_working(m, space);
        std::memcpy(out, f.data(), sizeof(float) * 9);
    });
}

rawops_status rawops_working_to_srgb_matrix(uint32_t space, float out[9]) {
    return guarded([&] {
        const auto f = working_to_srgb(space);
        std::memcpy(out, f.data(), sizeof(float) * 9);
    });
}

rawops_status rawops_working_to_working_matrix(uint32_t source, uint32_t target, float out[9]) {
    return guarded([&] {
        const auto f = working_to_working(source, target);
        std::memcpy(out, f.data(), sizeof(float) * 9);
    });
}

rawops_status rawops_xyz_d50_to_working_matrix(uint32_t space, double out[9]) {
    return guarded([&] {
        if (!out) fail(RAWOPS_INVALID_ARGUMENT, "null matrix pointer");
        const auto m = xyz_d50_to_working(space);
        for (int i = 0; i < 9; ++i) out[i] = m[i];
    });
}

rawops_status rawops_exposure_multiplier(float stops, float* out) {
    return guarded([&] {
        if (!std::isfinite(stops) || stops < -32.0f || stops > 32.0f)
            fail(RAWOPS_INVALID_ARGUMENT, "exposure stops must be finite and in [-32, 32]");
        *out = std::exp2(stops);
    });
}

rawops_status rawops_source_create_raw(const rawops_raw_metadata* meta, const uint16_t* host, size_t count,
                                       rawops_source** out) {
    return guarded([&] {
        if (!out) fail(RAWOPS_INVALID_ARGUMENT, "output handle is null");
        auto src = std::make_unique<rawops_source>();
        src->meta = normalise(meta, count);
        if (!host && count) fail(RAWOPS_INVALID_ARGUMENT, "samples are required");
        void* dev = nullptr;
        hip_check(hipMalloc(&dev, count * sizeof(uint16_t)), "hipMalloc(samples)");
        src->owned = dev;
        hip_check(hipMemcpy(dev, host, count * sizeof(uint16_t), hipMemcpyHostToDevice), "hipMemcpy(samples)");
        src->samples = static_cast<const uint16_t*>(dev);
        *out = src.release();
    });
}

rawops_status rawops_source_wrap_raw_device(const rawops_raw_metadata* meta, const uint16_t* dev, size_t count,
                                            rawops_source** out) {
    return guarded([&] {
        if (!out) fail(RAWOPS_INVALID_ARGUMENT, "output handle is null");
        auto src = std::make_unique<rawops_source>();
        src->meta = normalise(meta, count);
        if (!dev) fail(RAWOPS_INVALID_ARGUMENT, "samples are required");
        src->samples = dev;
        *out = src.release();
    });
}

static rawops_status raster_common(uint32_t w, uint32_t h, uint32_t stride, uint32_t space, const float* px,
                                   bool device, rawops_source** out) {
    return guarded([&] {
        if (!out) fail(RAWOPS_INVALID_ARGUMENT, "output handle is null");
        if (!w || !h || space > 1) fail(RAWOPS_INVALID_ARGUMENT, "invalid scene-linear raster metadata");
        if (!stride) stride = w;
        if (stride < w) fail(RAWOPS_INVALID_ARGUMENT, "raster float count does not match stride and height");
        if (!px) fail(RAWOPS_INVALID_ARGUMENT, "raster pixels are required");
        auto src = std::make_unique<rawops_source>();
        src->is_raw = false;
        src->width = w; src->height = h; src->stride = stride; src->space = space;
        if (device) {
            src->pixels = px;
        } else {
            const size_t n = size_t(stride) * h * 3;
            for (size_t i = 0; i < n; ++i)
                if (!std::isfinite(px[i])) fail(RAWOPS_INVALID_ARGUMENT, "scene-linear raster pixels must be finite");
            void* dev = nullptr;
            hip_check(hipMalloc(&dev, n * sizeof(float)), "hipMalloc(raster)");
            src->owned = dev;
            hip_check(hipMemcpy(dev, px, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(raster)");
            src->pixels = static_cast<const float*>(dev);
        }
        *out = src.release();
    });
}

rawops_status rawops_source_create_raster(uint32_t w, uint32_t h, uint32_t stride, uint32_t space,
                                          const float* host, rawops_source** out) {
    return raster_common(w, h, stride, space, host, false, out);
}

rawops_status rawops_source_wrap_raster_device(uint32_t w, uint32_t h, uint32_t stride, uint32_t space,
                                               const float* dev, rawops_source** out) {
    return raster_common(w, h, stride, space, dev, true, out);
}

rawops_status rawops_source_raw_metadata(const rawops_source* src, rawops_raw_metadata* out) {
    return guarded([&] {
        if (!src || !out || !src->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "RAW source required");
        *out = src->meta;
    });
}

void rawops_source_destroy(rawops_source* src) {
    if (!src) return;
    if (src->owned) (void)hipFree(src->owned);
    delete src;
}

rawops_status rawops_plan_create(const rawops_plan_desc* desc, rawops_plan** out) {
    return guarded([&] {
        if (!desc || !out) fail(RAWOPS_INVALID_ARGUMENT, "plan descriptor is required");
        if (desc->demosaic > RAWOPS_DEMOSAIC_MENON) fail(RAWOPS_INVALID_ARGUMENT, "unsupported RAW demosaic algorithm");
        auto plan = std::make_unique<rawops_plan>();
        plan->demosaic = desc->demosaic;
        plan->mip_anchor = desc->mip_anchor ? 1u : 0u;
        compile_ops(*plan, desc->native_ops, desc->n_native, plan->native);
        compile_ops(*plan, desc->reduced_ops, desc->n_reduced, plan->reduced);
        const OpList lists[3] = {plan->native, plan->reduced, OpList{}};
        hip_check(hipMalloc(reinterpret_cast<void**>(&plan->device), sizeof(lists)), "hipMalloc(plan)");
        hip_check(hipMemcpy(plan->device, lists, sizeof(lists), hipMemcpyHostToDevice), "hipMemcpy(plan)");
        *out = plan.release();
    });
}

void rawops_plan_destroy(rawops_plan* plan) { delete plan; }

rawops_status rawops_request_bounds(rawops_rect source, uint32_t mip, uint32_t quality, rawops_rect* out) {
    return guarded([&] { *out = level_bounds(source, mip, quality); });
}

rawops_status rawops_render(const rawops_plan* plan, const rawops_source* src, rawops_rect viewport,
                            uint32_t mip, uint32_t quality, float* dst, uint8_t* dst8, uint32_t dst_stride,
                            uint32_t* flags, void* stream) {
    return guarded([&] {
        validate_render(plan, src, viewport, mip, quality);
        if (!dst_stride) dst_stride = viewport.width;
        if (dst_stride < viewport.width) fail(RAWOPS_INVALID_ARGUMENT, "output stride is smaller than the viewport");
        if (!viewport.width || !viewport.height) return;
        RenderLaunch L{};
        L.sensor = sensor_args(*src);
        L.demosaic = plan->demosaic;
        // A singleton active-area axis routes Menon to bilinear (RawEngine.cpp:303).
        if (src->meta.active_area.width <= 1 || src->meta.active_area.height <= 1) L.demosaic = RAWOPS_DEMOSAIC_BILINEAR;
        L.mip = mip;
        L.viewport = viewport;
        L.out = dst;
        L.out8 = dst8;
        L.ostride = dst_stride;
        L.flags = flags;
        L.native_host = &plan->native;
        L.reduced_host = &plan->reduced;
        hip_check(launch_render(L, plan->device, plan->device + 1, static_cast<hipStream_t>(stream)), "render launch");
    });
}

rawops_status rawops_set_menon_path(uint32_t path, uint32_t band_rows) {
    return guarded([&] {
        if (path > 3)
            fail(RAWOPS_INVALID_ARGUMENT,
                 "Menon render path must be 0 (automatic), 1 (streaming), 2 (tiles) or 3 (streaming, FP32 stages)");
        if (band_rows && (band_rows < 4 || band_rows > 4096 || band_rows % 4))
            fail(RAWOPS_INVALID_ARGUMENT, "streaming band rows must be 0 or a multiple of 4 in 4..4096");
        g_menon_path.store(path, std::memory_order_relaxed);
        g_stream_band.store(band_rows, std::memory_order_relaxed);
    });
}

rawops_status rawops_menon_decisions_size(const rawops_source* src, size_t* bytes) {
    return guarded([&] {
        if (!src || !src->is_raw || !bytes) fail(RAWOPS_INVALID_ARGUMENT, "decisions need a RAW source");
        *bytes = decision_bytes(src->meta.active_area.width, src->meta.active_area.height);
    });
}

rawops_status rawops_render_record(const rawops_plan* plan, const rawops_source* src, rawops_rect viewport,
                                   float* dst, uint32_t dst_stride, uint32_t* flags, uint8_t* decisions,
                                   void* stream) {
    return guarded([&] {
        validate_render(plan, src, viewport, 0, RAWOPS_QUALITY_FINAL);
        if (plan->demosaic != RAWOPS_DEMOSAIC_MENON) fail(RAWOPS_INVALID_ARGUMENT, "decisions are recorded by Menon renders");
        if (!decisions) fail(RAWOPS_INVALID_ARGUMENT, "decision plane is required");
        if (reinterpret_cast<uintptr_t>(decisions) & 3) fail(RAWOPS_INVALID_ARGUMENT, "decision plane must be 4-byte aligned");
        if (!dst_stride) dst_stride = viewport.width;
        if (dst_stride < viewport.width) fail(RAWOPS_INVALID_ARGUMENT, "output stride is smaller than the viewport");
        if (!viewport.width || !viewport.height) return;
        if (src->meta.active_area.width <= 1 || src->meta.active_area.height <= 1)
            fail(RAWOPS_INVALID_ARGUMENT, "a singleton active-area axis renders bilinear (no decisions)");
        RenderLaunch L{};
        L.sensor = sensor_args(*src);
        L.demosaic = RAWOPS_DEMOSAIC_MENON;
        L.mip = 0;
        L.viewport = viewport;
        L.out = dst;
        L.ostride = dst_stride;
        L.flags = flags;
        L.decisions = reinterpret_cast<uint32_t*>(decisions);
        L.native_host = &plan->native;
        L.reduced_host = &plan->reduced;
        hip_check(launch_render(L, plan->device, plan->device + 1, static_cast<hipStream_t>(stream)), "render launch");
    });
}

rawops_status rawops_check_flags(uint32_t* flags, void* stream) {
    return guarded([&] {
        if (!flags) return;
        uint32_t host = 0;
        auto s = static_cast<hipStream_t>(stream);
        hip_check(hipMemcpyAsync(&host, flags, sizeof(host), hipMemcpyDeviceToHost, s), "flag readback");
        hip_check(hipStreamSynchronize(s), "stream synchronize");
        if (host) hip_check(hipMemsetAsync(flags, 0, sizeof(uint32_t), s), "flag reset");
        if (host & FLAG_OVERFLOW) fail(RAWOPS_OVERFLOW_ERROR, "guided filter nonfinite intermediate");
        if (host & FLAG_INVALID) fail(RAWOPS_INVALID_ARGUMENT, "extended curves/levels output exceeds finite float32");
        if (host & FLAG_DOMAIN) fail(RAWOPS_DOMAIN_ERROR, "non-finite value at output boundary");
    });
}

rawops_status rawops_render_tiles_host(const rawops_plan* plan, const rawops_source* src, rawops_rect viewport,
                                       uint32_t mip, uint32_t quality, uint32_t tile_size,
                                       rawops_tile_callback callback, void* user, const volatile int* cancel) {
    return guarded([&] {
        validate_render(plan, src, viewport, mip, quality);
        if (!callback || !tile_size) fail(RAWOPS_INVALID_ARGUMENT, "callback and tile size are required");
        if (cancel && *cancel) fail(RAWOPS_CANCELLED, "render cancelled");
        if (!viewport.width || !viewport.height) return;
        const size_t n = size_t(viewport.width) * viewport.height * 3;
        float* dev = nullptr;
        uint32_t* flags = nullptr;
        float* host = nullptr;
        struct Cleanup {
            float*& d; uint32_t*& f; float*& h;
            ~Cleanup() { if (d) (void)hipFree(d); if (f) (void)hipFree(f); if (h) (void)hipHostFree(h); }
        } cleanup{dev, flags, host};
        hip_check(hipMalloc(&dev, n * sizeof(float)), "hipMalloc(viewport)");
        hip_check(hipMalloc(&flags, sizeof(uint32_t)), "hipMalloc(flags)");
        hip_check(hipHostMalloc(&host, n * sizeof(float), hipHostMallocDefault), "hipHostMalloc(viewport)");
        hip_check(hipMemset(flags, 0, sizeof(uint32_t)), "hipMemset(flags)");
        const rawops_status st = rawops_render(plan, src, viewport, mip, quality, dev, nullptr, viewport.width,
                                               flags, nullptr);
        if (st != RAWOPS_OK) fail(st, g_error);
        hip_check(hipMemcpy(host, dev, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(viewport)");
        const rawops_status fs = rawops_check_flags(flags, nullptr);
        if (fs != RAWOPS_OK) fail(fs, g_error);
        std::vector<float> tile;
        const uint64_t right = uint64_t(viewport.x) + viewport.width, bottom = uint64_t(viewport.y) + viewport.height;
        for (uint64_t y = viewport.y; y < bottom; y += tile_size)
            for (uint64_t x = viewport.x; x < right; x += tile_size) {
                if (cancel && *cancel) fail(RAWOPS_CANCELLED, "render cancelled");
                const rawops_rect r{uint32_t(x), uint32_t(y), uint32_t(std::min<uint64_t>(tile_size, right - x)),
                                    uint32_t(std::min<uint64_t>(tile_size, bottom - y))};
                tile.resize(size_t(r.width) * r.height * 3);
                for (uint32_t row = 0; row < r.height; ++row)
                    std::memcpy(tile.data() + size_t(row) * r.width * 3,
                                host + ((size_t(r.y - viewport.y) + row) * viewport.width + (r.x - viewport.x)) * 3,
                                size_t(r.width) * 3 * sizeof(float));
                if (callback(user, r, tile.data())) fail(RAWOPS_CANCELLED, "render cancelled by callback");
            }
    });
}

static void batch_layout(rawops_batch& b, int shape) {
    const auto sh = kBatchShapes[shape];
    int first = 0;
    for (auto& it : b.items) {
        const int scale = 1 << it.mip;
        const int lw = sh.tw / scale, lh = sh.th / scale;
        it.tiles_x = (it.vw + lw - 1) / lw;
        it.tiles = it.tiles_x * ((it.vh + lh - 1) / lh);
        it.first = first;
        first += it.tiles;
    }
    b.total = first;
    b.shape = shape;
    hip_check(hipMemcpy(b.dev_items, b.items.data(), b.items.size() * sizeof(BatchItem), hipMemcpyHostToDevice),
              "hipMemcpy(batch)");
}

rawops_status rawops_batch_create(const rawops_batch_request* reqs, size_t count, uint32_t shape_index, void* stream,
                                  rawops_batch** out) {
    return guarded([&] {
        if (!out || (!reqs && count)) fail(RAWOPS_INVALID_ARGUMENT, "batch requests are required");
        if (!count) fail(RAWOPS_INVALID_ARGUMENT, "a batch needs at least one request");
        if (shape_index != UINT32_MAX && shape_index >= uint32_t(kBatchShapeCount))
            fail(RAWOPS_INVALID_ARGUMENT, "unknown batch tile shape");
        auto b = std::make_unique<rawops_batch>();
        b->order.resize(count);
        std::iota(b->order.begin(), b->order.end(), size_t(0));
        std::stable_sort(b->order.begin(), b->order.end(), [&](size_t x, size_t y) {
            return uint64_t(reqs[x].viewport.width) * reqs[x].viewport.height <
                   uint64_t(reqs[y].viewport.width) * reqs[y].viewport.height;
        });
        hip_check(hipMalloc(reinterpret_cast<void**>(&b->flags), sizeof(uint32_t)), "hipMalloc(flags)");
        hip_check(hipMemset(b->flags, 0, sizeof(uint32_t)), "hipMemset(flags)");
        for (size_t k = 0; k < count; ++k) {
            const rawops_batch_request& r = reqs[b->order[k]];
            validate_render(r.plan, r.source, r.viewport, r.mip, r.quality);
            if (!r.viewport.width || !r.viewport.height) fail(RAWOPS_INVALID_ARGUMENT, "batch viewports must be nonempty");
            BatchItem it{};
            it.sensor = sensor_args(*r.source);
            it.native = r.plan->device;
            it.reduced = r.plan->device + 1;
            it.tap = r.linear_ops;
            it.demosaic = int(r.plan->demosaic);
            if (r.source->meta.active_area.width <= 1 || r.source->meta.active_area.height <= 1)
                it.demosaic = RAWOPS_DEMOSAIC_BILINEAR;
            it.mip = int(r.mip);
            it.vx = int(r.viewport.x); it.vy = int(r.viewport.y);
            it.vw = int(r.viewport.width); it.vh = int(r.viewport.height);
            const int scale = 1 << it.mip;
            it.ox0 = it.mip ? it.sensor.ax + it.vx * scale : it.vx;
            it.oy0 = it.mip ? it.sensor.ay + it.vy * scale : it.vy;
            it.out = r.dst_rgb;
            it.out8 = r.dst_u8;
            it.ostride = r.dst_stride_pixels ? r.dst_stride_pixels : r.viewport.width;
            if (it.ostride < r.viewport.width) fail(RAWOPS_INVALID_ARGUMENT, "output stride is smaller than the viewport");
            it.flags = b->flags;
            b->items.push_back(it);
        }
        const size_t n = b->items.size();
        hip_check(hipMalloc(reinterpret_cast<void**>(&b->dev_items), n * sizeof(BatchItem)), "hipMalloc(items)");
        hip_check(hipMalloc(&b->state, batch_state_bytes()), "hipMalloc(state)");
        hip_check(hipMalloc(reinterpret_cast<void**>(&b->remaining), n * sizeof(uint32_t)), "hipMalloc(remaining)");
        hip_check(hipMalloc(reinterpret_cast<void**>(&b->done), n * sizeof(unsigned long long)), "hipMalloc(done)");
        auto st = static_cast<hipStream_t>(stream);
        int shape = shape_index == UINT32_MAX ? 0 : int(shape_index);
        if (shape_index == UINT32_MAX) {
            // Autotune: time every compiled tile shape on this batch.
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            double best = 1e30;
            for (int c = 0; c < kBatchShapeCount; ++c) {
                batch_layout(*b, c);
                hip_check(launch_batch(c, b->dev_items, int(n), b->total, b->state, b->remaining, b->done, st),
                          "batch launch");
                hip_check(hipEventRecord(e0, st), "hipEventRecord");
                for (int rep = 0; rep < 3; ++rep)
                    hip_check(launch_batch(c, b->dev_items, int(n), b->total, b->state, b->remaining, b->done, st),
                              "batch launch");
                hip_check(hipEventRecord(e1, st), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                if (ms / 3.0 < best) { best = ms / 3.0; shape = c; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            b->tuned_ms = best;
        }
        batch_layout(*b, shape);
        // Capture on a private stream (the legacy null stream cannot be
        // captured); the graph is replayed on the caller's stream.
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        hipStream_t cap = nullptr;
        hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        struct Destroy { hipStream_t s; ~Destroy() { (void)hipStreamDestroy(s); } } destroy{cap};
        hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const hipError_t le = launch_batch(shape, b->dev_items, int(n), b->total, b->state, b->remaining, b->done,
                                           cap);
        const hipError_t ce = hipStreamEndCapture(cap, &b->graph);
        hip_check(le, "batch launch (capture)");
        hip_check(ce, "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        *out = b.release();
    });
}

rawops_status rawops_batch_run(rawops_batch* b, void* stream) {
    return guarded([&] {
        if (!b) fail(RAWOPS_INVALID_ARGUMENT, "batch is null");
        hip_check(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
    });
}

rawops_status rawops_batch_latencies(rawops_batch* b, double* ms) {
    return guarded([&] {
        if (!b || !ms) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        const size_t n = b->items.size();
        std::vector<unsigned long long> done(n);
        std::vector<unsigned char> state(batch_state_bytes());
        hip_check(hipMemcpy(done.data(), b->done, n * sizeof(unsigned long long), hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(state.data(), b->state, state.size(), hipMemcpyDeviceToHost), "hipMemcpy");
        const unsigned long long t0 = batch_start_time(state.data());
        int dev = 0, khz = 100000;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
        for (size_t k = 0; k < n; ++k) ms[b->order[k]] = double(done[k] - t0) / double(khz);
    });
}

rawops_status rawops_batch_check(rawops_batch* b, void* stream) {
    return guarded([&] {
        if (!b) fail(RAWOPS_INVALID_ARGUMENT, "batch is null");
        const rawops_status st = rawops_check_flags(b->flags, stream);
        if (st != RAWOPS_OK) fail(st, g_error);
    });
}

rawops_status rawops_batch_shape(const rawops_batch* b, uint32_t* tw, uint32_t* th, double* tuned_ms) {
    return guarded([&] {
        if (!b) fail(RAWOPS_INVALID_ARGUMENT, "batch is null");
        if (tw) *tw = uint32_t(kBatchShapes[b->shape].tw);
        if (th) *th = uint32_t(kBatchShapes[b->shape].th);
        if (tuned_ms) *tuned_ms = b->tuned_ms;
    });
}

uint32_t rawops_batch_shape_count(void) { return uint32_t(kBatchShapeCount); }

void rawops_batch_destroy(rawops_batch* b) { delete b; }

rawops_status rawops_render_backward(const rawops_source* src, uint32_t demosaic, rawops_rect viewport,
                                     const float gains[3], float multiplier, const float matrix[9],
                                     const float* g_out, uint32_t g_stride, float* g_raw, double* g_params,
                                     void* stream) {
    return guarded([&] {
        if (!src || !src->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "backward needs a RAW source");
        if (demosaic > RAWOPS_DEMOSAIC_MENON) fail(RAWOPS_INVALID_ARGUMENT, "unsupported RAW demosaic algorithm");
        if (!gains || !matrix || !g_out || !g_params) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        inside_or_throw(src->meta.active_area, viewport);
        if (!g_stride) g_stride = viewport.width;
        BackwardLaunch L{};
        L.sensor = sensor_args(*src);
        L.demosaic = demosaic;
        if (src->meta.active_area.width <= 1 || src->meta.active_area.height <= 1) L.demosaic = RAWOPS_DEMOSAIC_BILINEAR;
        L.viewport = viewport;
        for (int i = 0; i < 3; ++i) L.gains[i] = gains[i];
        L.multiplier = multiplier;
        for (int i = 0; i < 9; ++i) L.matrix[i] = matrix[i];
        L.g_out = g_out;
        L.g_stride = g_stride;
        L.g_raw = g_raw;
        L.g_params = g_params;
        hip_check(launch_render_backward(L, static_cast<hipStream_t>(stream)), "backward launch");
    });
}

rawops_status rawops_render_backward_recorded(const rawops_source* src, rawops_rect viewport, const float gains[3],
                                              float multiplier, const float matrix[9], const float* g_out,
                                              uint32_t g_stride, float* g_raw, double* g_params,
                                              const uint8_t* decisions, rawops_rect recorded, void* stream) {
    return guarded([&] {
        if (!src || !src->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "backward needs a RAW source");
        if (!gains || !matrix || !g_out || !g_params || !decisions) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        if (reinterpret_cast<uintptr_t>(decisions) & 3) fail(RAWOPS_INVALID_ARGUMENT, "decision plane must be 4-byte aligned");
        inside_or_throw(src->meta.active_area, viewport);
        inside_or_throw(src->meta.active_area, recorded);
        if (src->meta.active_area.width <= 1 || src->meta.active_area.height <= 1)
            fail(RAWOPS_INVALID_ARGUMENT, "a singleton active-area axis renders bilinear (no decisions)");
        if (!g_stride) g_stride = viewport.width;
        BackwardLaunch L{};
        L.sensor = sensor_args(*src);
        L.demosaic = RAWOPS_DEMOSAIC_MENON;
        L.viewport = viewport;
        for (int i = 0; i < 3; ++i) L.gains[i] = gains[i];
        L.multiplier = multiplier;
        for (int i = 0; i < 9; ++i) L.matrix[i] = matrix[i];
        L.g_out = g_out;
        L.g_stride = g_stride;
        L.g_raw = g_raw;
        L.g_params = g_params;
        L.decisions = reinterpret_cast<const uint32_t*>(decisions);
        L.recorded = recorded;
        hip_check(launch_render_backward(L, static_cast<hipStream_t>(stream)), "backward launch");
    });
}

rawops_status rawops_srgb_encode(const float* x, float* y, size_t n, void* stream) {
    return guarded([&] { hip_check(launch_srgb(x, y, n, static_cast<hipStream_t>(stream)), "srgb launch"); });
}

rawops_status rawops_srgb_encode_backward(const float* x, const float* gy, float* gx, size_t n, void* stream) {
    return guarded([&] {
        hip_check(launch_srgb_backward(x, gy, gx, n, static_cast<hipStream_t>(stream)), "srgb backward launch");
    });
}

rawops_status rawops_exposure_backward(const float* x, const float* gy, float m, float* gx, double* g_stops, size_t n,
                                       void* stream) {
    return guarded([&] {
        hip_check(launch_exposure_backward(x, gy, m, gx, g_stops, n, static_cast<hipStream_t>(stream)),
                  "exposure backward launch");
    });
}

rawops_status rawops_detail_boost(const float* x, const float* base, float strength, float* y, size_t n, void* stream) {
    return guarded([&] {
        hip_check(launch_boost(x, base, strength, y, n, static_cast<hipStream_t>(stream)), "boost launch");
    });
}

rawops_status rawops_detail_boost_backward(const float* x, const float* base, const float* gy, float strength,
                                           float* gx, float* gbase, double* gs, size_t n, void* stream) {
    return guarded([&] {
        hip_check(launch_boost_backward(x, base, gy, strength, gx, gbase, gs, n, static_cast<hipStream_t>(stream)),
                  "boost backward launch");
    });
}

rawops_status rawops_curve_backward(const rawops_plan* plan, const float* x, const float* gy, size_t n, float* gx,
                                    double* gys, void* stream) {
    return guarded([&] {
        if (!plan || plan->reduced.n < 1 || plan->reduced.op[0].kind != OP_CURVES)
            fail(RAWOPS_INVALID_ARGUMENT, "curve backward needs a plan whose first reduced op is a curve");
        const DevCurves* t = plan->reduced.op[0].curves;
        hip_check(launch_curve_backward(x, gy, n, &t->c[0], plan->reduced.op[0].knots, plan->curve_cubic, gx, gys,
                                        static_cast<hipStream_t>(stream)),
                  "curve backward launch");
    });
}

rawops_status rawops_exposure_curve_backward(const rawops_plan* plan, const float* x, float m, const float* gy,
                                             size_t n, float* gx, double* gys, double* g_stops, void* stream) {
    return guarded([&] {
        if (!plan || plan->reduced.n < 1 || plan->reduced.op[0].kind != OP_CURVES)
            fail(RAWOPS_INVALID_ARGUMENT, "curve backward needs a plan whose first reduced op is a curve");
        if (!g_stops) fail(RAWOPS_INVALID_ARGUMENT, "g_stops is null");
        const DevCurves* t = plan->reduced.op[0].curves;
        hip_check(launch_curve_backward(x, gy, n, &t->c[0], plan->reduced.op[0].knots, plan->curve_cubic, gx, gys,
                                        static_cast<hipStream_t>(stream), m, g_stops),
                  "exposure + curve backward launch");
    });
}

rawops_status rawops_guided_backward_workspace_size(uint32_t width, uint32_t height, size_t* bytes) {
    return guarded([&] {
        if (!bytes) fail(RAWOPS_INVALID_ARGUMENT, "output pointer is null");
        *bytes = guided_backward_workspace_bytes(width, height);
    });
}

rawops_status rawops_guided_backward(const rawops_source* raster, uint32_t radius, double epsilon, rawops_rect viewport,
                                     const float* g_out, uint32_t g_stride, float* g_in, double* g_eps,
                                     void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        if (!raster || raster->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires a scene-linear raster");
        if (radius < 1 || radius > 32 || !std::isfinite(epsilon) || epsilon < 0x1p-24 || epsilon > 65536.0)
            fail(RAWOPS_INVALID_ARGUMENT, "guided backward requires radius 1..32 and finite epsilon [2^-24,65536]");
        inside_or_throw({0, 0, raster->width, raster->height}, viewport);
        if (!g_out || !g_in || !g_eps) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        if (!g_stride) g_stride = viewport.width;
        GuidedBackwardLaunch L{};
        L.src = raster->pixels;
        L.src_stride = raster->stride;
        L.width = raster->width;
        L.height = raster->height;
        L.radius = radius;
        L.epsilon = epsilon;
        const auto w = working_y_weights(raster->space);
        L.wr = w[0];
        L.wb = w[1];
        L.g_out = g_out;
        L.g_stride = g_stride;
        L.viewport = viewport;
        L.g_in = g_in;
        L.g_eps = g_eps;
        auto st = static_cast<hipStream_t>(stream);
        const size_t need = guided_backward_workspace_bytes(raster->width, raster->height);
        void* owned = nullptr;
        if (workspace) {
            if (workspace_bytes < need) fail(RAWOPS_INVALID_ARGUMENT, "guided backward workspace is too small");
            L.workspace = workspace;
        } else {
            hip_check(hipMallocAsync(&owned, need, st), "hipMallocAsync(guided backward workspace)");
            L.workspace = owned;
        }
        const hipError_t e = launch_guided_backward(L, st);
        if (owned) (void)hipFreeAsync(owned, st);
        hip_check(e, "guided backward launch");
    });
}

rawops_status rawops_histogram_rgb(const float* rgb, uint32_t width, uint32_t height, uint32_t stride_pixels,
                                   uint32_t bins, double lower, double upper, uint64_t* counts, uint64_t* tallies,
                                   uint32_t* extrema, void* stream) {
    return guarded([&] {
        const double span = upper - lower;
        if (bins == 0 || bins > 65536 || !std::isfinite(lower) || !std::isfinite(upper) || !(span > 0.0) ||
            !std::isfinite(span))
            fail(RAWOPS_INVALID_ARGUMENT, "histogram requires 1..65536 bins and a finite positive range");
        if (!counts || !tallies || !extrema) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        if (!stride_pixels) stride_pixels = width;
        if (stride_pixels < width) fail(RAWOPS_INVALID_ARGUMENT, "image stride is smaller than its width");
        if (!width || !height) return;
        if (!rgb) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        HistogramLaunch h{};
        h.rgb = rgb;
        h.width = width;
        h.height = height;
        h.stride = stride_pixels;
        h.bins = bins;
        h.lower = lower;
        h.upper = upper;
        h.span = span;
        h.counts = reinterpret_cast<unsigned long long*>(counts);
        h.tallies = reinterpret_cast<unsigned long long*>(tallies);
        h.extrema = extrema;
        hip_check(launch_histogram(h, static_cast<hipStream_t>(stream)), "histogram launch");
    });
}

rawops_status rawops_local_statistics(const float* rgb, rawops_rect needed, uint32_t stride_pixels,
                                      rawops_rect output, rawops_rect image, uint32_t radius, double* mean,
                                      double* variance, uint32_t* finite_count, void* stream) {
    return guarded([&] {
        if (radius > 8) fail(RAWOPS_INVALID_ARGUMENT, "local statistics radius must be in 0..8");
        auto addressable = [](rawops_rect r) {
            return r.width && r.height && uint64_t(r.x) + r.width <= (uint64_t{1} << 32) &&
                   uint64_t(r.y) + r.height <= (uint64_t{1} << 32);
        };
        if (!addressable(output) || !addressable(image))
            fail(RAWOPS_INVALID_ARGUMENT, "spatial operation requires nonempty addressable bounds");
        if (output.x < image.x || output.y < image.y ||
            uint64_t(output.x) + output.width > uint64_t(image.x) + image.width ||
            uint64_t(output.y) + output.height > uint64_t(image.y) + image.height)
            fail(RAWOPS_OUT_OF_RANGE, "spatial viewport is outside image bounds");
        const uint64_t left = std::max<int64_t>(image.x, int64_t(output.x) - radius);
        const uint64_t top = std::max<int64_t>(image.y, int64_t(output.y) - radius);
        const uint64_t right = std::min<uint64_t>(uint64_t(image.x) + image.width, uint64_t(output.x) + output.width + radius);
        const uint64_t bottom = std::min<uint64_t>(uint64_t(image.y) + image.height, uint64_t(output.y) + output.height + radius);
        if (needed.x != left || needed.y != top || needed.width != right - left || needed.height != bottom - top)
            fail(RAWOPS_INVALID_ARGUMENT, "local statistics input must cover exactly the output grown by the radius");
        if (!stride_pixels) stride_pixels = needed.width;
        if (stride_pixels < needed.width) fail(RAWOPS_INVALID_ARGUMENT, "image stride is smaller than its width");
        if (!rgb || !mean || !variance || !finite_count) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        LocalStatsLaunch L{};
        L.in = rgb;
        L.in_stride = stride_pixels;
        L.needed = needed;
        L.output = output;
        L.image = image;
        L.radius = radius;
        L.mean = mean;
        L.variance = variance;
        L.count = finite_count;
        hip_check(launch_local_stats(L, static_cast<hipStream_t>(stream)), "local statistics launch");
    });
}

rawops_status rawops_guided_filter_stack(const rawops_source* raster, uint32_t images, rawops_rect viewport,
                                         uint32_t radius, double epsilon, const rawops_plan* post, float* dst,
                                         uint32_t dst_stride, uint32_t* flags, void* stream) {
    return guarded([&] {
        if (!raster || raster->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires a scene-linear raster");
        if (!images || raster->height % images) fail(RAWOPS_INVALID_ARGUMENT, "the raster height must split into the images");
        if (radius > 32 || !std::isfinite(epsilon) || epsilon < 0x1p-24 || epsilon > 65536.0)
            fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires radius 0..32 and finite epsilon [2^-24,65536]");
        if (!post || !dst) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        const uint32_t H = raster->height / images;
        inside_or_throw({0, 0, raster->width, H}, viewport);
        if (!dst_stride) dst_stride = viewport.width;
        if (dst_stride < viewport.width) fail(RAWOPS_INVALID_ARGUMENT, "output stride is smaller than the viewport");
        if (!viewport.width || !viewport.height) return;
        GuidedLaunch L{};
        L.src = raster->pixels;
        L.src_stride = raster->stride;
        L.width = raster->width;
        L.height = H;
        L.mip = 0;
        L.viewport = viewport;
        L.radius = radius;
        L.epsilon = epsilon;
        const auto w = working_y_weights(raster->space);
        L.wr = w[0];
        L.wb = w[1];
        L.out = dst;
        L.ostride = dst_stride;
        L.flags = flags;
        L.slices = images;
        L.src_zstride = (size_t)H * raster->stride * 3;
        L.out_zstride = (size_t)viewport.height * dst_stride * 3;
        L.post_ops = L.fused_ops = post->reduced.n;
        for (uint32_t i = 0; i < post->reduced.n; ++i)
            if (post->reduced.op[i].kind == OP_CURVES) { L.fused_ops = i; break; }
        auto st = static_cast<hipStream_t>(stream);
        const size_t need = radius ? guided_workspace_bytes(L) : 0;
        void* owned = nullptr;
        if (need) {
            hip_check(hipMallocAsync(&owned, need, st), "hipMallocAsync(guided workspace)");
            L.workspace = owned;
        }
        const hipError_t e = launch_guided(L, post->device + 1, st);
        if (owned) (void)hipFreeAsync(owned, st);
        hip_check(e, "guided launch");
    });
}

rawops_status rawops_guided_backward_stack(const rawops_source* raster, uint32_t images, uint32_t radius,
                                           double epsilon, rawops_rect viewport, const float* g_out,
                                           uint32_t g_stride, float* g_in, double* g_eps, void* stream) {
    return guarded([&] {
        if (!raster || raster->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires a scene-linear raster");
        if (!images || raster->height % images) fail(RAWOPS_INVALID_ARGUMENT, "the raster height must split into the images");
        if (radius < 1 || radius > 32 || !std::isfinite(epsilon) || epsilon < 0x1p-24 || epsilon > 65536.0)
            fail(RAWOPS_INVALID_ARGUMENT, "guided backward requires radius 1..32 and finite epsilon [2^-24,65536]");
        const uint32_t H = raster->height / images;
        inside_or_throw({0, 0, raster->width, H}, viewport);
        if (!g_out || !g_in || !g_eps) fail(RAWOPS_INVALID_ARGUMENT, "null argument");
        if (!g_stride) g_stride = viewport.width;
        GuidedBackwardLaunch L{};
        L.src = raster->pixels;
        L.src_stride = raster->stride;
        L.width = raster->width;
        L.height = H;
        L.radius = radius;
        L.epsilon = epsilon;
        const auto w = working_y_weights(raster->space);
        L.wr = w[0];
        L.wb = w[1];
        L.g_out = g_out;
        L.g_stride = g_stride;
        L.viewport = viewport;
        L.g_in = g_in;
        L.g_eps = g_eps;
        L.slices = images;
        L.src_zstride = (size_t)H * raster->stride * 3;
        L.g_out_zstride = (size_t)viewport.height * g_stride * 3;
        auto st = static_cast<hipStream_t>(stream);
        void* owned = nullptr;
        hip_check(hipMallocAsync(&owned, guided_backward_workspace_bytes(raster->width, H) * images, st),
                  "hipMallocAsync(guided backward workspace)");
        L.workspace = owned;
        const hipError_t e = launch_guided_backward(L, st);
        (void)hipFreeAsync(owned, st);
        hip_check(e, "guided backward launch");
    });
}

rawops_status rawops_guided_workspace_size(uint32_t width, uint32_t height, rawops_rect viewport, uint32_t mip,
                                           uint32_t radius, size_t* bytes) {
    return guarded([&] {
        if (!bytes) fail(RAWOPS_INVALID_ARGUMENT, "output pointer is null");
        GuidedLaunch L{};
        L.width = width; L.height = height; L.viewport = viewport; L.mip = mip; L.radius = radius;
        *bytes = radius ? guided_workspace_bytes(L) : 0;
    });
}

}  // extern "C"

// Scratch of rawops_guided_filter_static: the per-centre coefficient planes
// (6 floats per centre) of one call.  One copy per device, held until the
// process ends; calls that overlap on one device share it and race.
static constexpr size_t kGuidedScratchBytes = size_t(RAWOPS_GUIDED_STATIC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kGuidedScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with the device current).  The array's name in
// host code is not a device address; only this pointer is passed on.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

void* guided_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) fail(RAWOPS_DEVICE_ERROR, "device ordinal beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) fail(RAWOPS_DEVICE_ERROR, "static scratch has no device address");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Shared body of rawops_guided_filter and rawops_guided_filter_static.
// workspace: caller memory of workspace_bytes, the static scratch (static_ws),
// or NULL (allocated stream-ordered).
void guided_filter_body(const rawops_source* raster, rawops_rect viewport, uint32_t mip, uint32_t radius,
                        double epsilon, const rawops_plan* post, float* dst, uint32_t dst_stride, void* workspace,
                        size_t workspace_bytes, bool static_ws, uint32_t* flags, void* stream) {
    if (!raster || raster->is_raw) fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires a scene-linear raster");
    if (radius > 32 || !std::isfinite(epsilon) || epsilon < 0x1p-24 || epsilon > 65536.0)
        fail(RAWOPS_INVALID_ARGUMENT, "guided filter requires radius 0..32 and finite epsilon [2^-24,65536]");
    const rawops_rect image = level_bounds({0, 0, raster->width, raster->height}, mip,
                                           mip ? RAWOPS_QUALITY_PREVIEW : RAWOPS_QUALITY_FINAL);
    if (viewport.x < image.x || viewport.y < image.y ||
        uint64_t(viewport.x) + viewport.width > uint64_t(image.x) + image.width ||
        uint64_t(viewport.y) + viewport.height > uint64_t(image.y) + image.height)
        fail(RAWOPS_OUT_OF_RANGE, "spatial viewport is outside image bounds");
    if (!dst_stride) dst_stride = viewport.width;
    if (!viewport.width || !viewport.height) return;
    GuidedLaunch L{};
    L.src = raster->pixels;
    L.src_stride = raster->stride;
    L.width = raster->width;
    L.height = raster->height;
    L.mip = mip;
    L.viewport = viewport;
    L.radius = radius;
    L.epsilon = epsilon;
    const auto w = working_y_weights(raster->space);
    L.wr = w[0];
    L.wb = w[1];
    L.out = dst;
    L.ostride = dst_stride;
    L.flags = flags;
    if (!post) fail(RAWOPS_INVALID_ARGUMENT, "guided filter needs a post-filter plan (may be empty)");
    auto st = static_cast<hipStream_t>(stream);
    const size_t need = radius ? guided_workspace_bytes(L) : 0;
    void* owned = nullptr;
    if (static_ws) {
        if (need > kGuidedScratchBytes)
            fail(RAWOPS_LENGTH_ERROR, "guided filter workspace of " + std::to_string(need) +
                                          " bytes exceeds the static scratch of " +
                                          std::to_string(kGuidedScratchBytes) + " bytes");
        if (need) L.workspace = guided_scratch();
    } else if (need) {
        if (workspace) {
            if (workspace_bytes < need) fail(RAWOPS_INVALID_ARGUMENT, "guided filter workspace is too small");
            L.workspace = workspace;
        } else {
            hip_check(hipMallocAsync(&owned, need, st), "hipMallocAsync(guided workspace)");
            L.workspace = owned;
        }
    }
    // Curve lookups are long-latency (data-dependent table reads, an FP64
    // division): inside the strip pass they stall its per-row loop, so
    // they run in a separate pointwise pass over the output.
    L.post_ops = L.fused_ops = post->reduced.n;
    for (uint32_t i = 0; i < post->reduced.n; ++i)
        if (post->reduced.op[i].kind == OP_CURVES) { L.fused_ops = i; break; }
    const hipError_t e = launch_guided(L, post->device + 1, st);
    if (owned) (void)hipFreeAsync(owned, st);
    hip_check(e, "guided launch");
}

}  // namespace

extern "C" {

rawops_status rawops_guided_filter(const rawops_source* raster, rawops_rect viewport, uint32_t mip,
                                   uint32_t radius, double epsilon, const rawops_plan* post, float* dst,
                                   uint32_t dst_stride, void* workspace, size_t workspace_bytes,
                                   uint32_t* flags, void* stream) {
    return guarded([&] {
        guided_filter_body(raster, viewport, mip, radius, epsilon, post, dst, dst_stride, workspace,
                           workspace_bytes, false, flags, stream);
    });
}

rawops_status rawops_guided_filter_static(const rawops_source* raster, rawops_rect viewport, uint32_t mip,
                                          uint32_t radius, double epsilon, const rawops_plan* post, float* dst,
                                          uint32_t dst_stride, uint32_t* flags, void* stream) {
    return guarded([&] {
        guided_filter_body(raster, viewport, mip, radius, epsilon, post, dst, dst_stride, nullptr, 0, true, flags,
                           stream);
    });
}

}  // extern "C"
