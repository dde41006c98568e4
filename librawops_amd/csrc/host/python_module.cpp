// _rawengine_native (imported as librawops_amd.rawengine_native): the
// reference's Python entry points for the render path
// (RawEnginePython.cpp:146-519 one-shot renders, 735-2256 sessions and jobs)
// over this library's C++ drop-in layer: same function and class names, option
// keys, result tuples and exception classes.  Renders run on the GPU; the GIL
// is released around native work.
//
// Extensions (no counterpart in the reference): `render_dlpack` (module and
// sessions) returns the result as a DLPack capsule of device memory
// (float32, shape (height, width, 3), on the ROCm device) instead of a copy
// into `bytes`: `torch.from_dlpack(capsule)` adopts it without a transfer.
//
// Not here (outside the render path): ICC profile creation (no colour engine
// in this build: icc_available() is False, as in a reference build without
// LittleCMS), raster crop/resize/rotation options, RasterGraphSession,
// coverage/RGBA outputs, EditHistory, Cube LUT import.
#include "../../../include/rawengine/EditGraph.hpp"
#include "../../../include/rawengine/ImageAnalysis.hpp"
#include "../../../include/rawengine/RawEngine.hpp"
#include "../../../include/rawengine/SpatialOps.hpp"
#include "../../../include/rawengine/TileScheduler.hpp"
#include "internal.hpp"

#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

namespace py = pybind11;
using namespace rawengine;

namespace {

PyObject* g_cancelled = nullptr;  // rawengine_native.RenderCancelled

class SessionClosed : public std::logic_error {
public:
    SessionClosed() : std::logic_error("render session is closed") {}
};

// Conditions the reference reports as RuntimeError (a full submission queue).
class QueueFull : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

[[noreturn]] void raise_py(PyObject* type, const char* what) {
    PyErr_SetString(type, what);
    throw py::error_already_set();
}

// Translate like the reference: invalid options and domain errors are
// ValueError, cancellation RenderCancelled, a closed session or a full queue
// RuntimeError.
template <class F> auto guarded(F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const py::error_already_set&) {
        throw;
    } catch (const py::cast_error& e) {
        raise_py(PyExc_TypeError, e.what());
    } catch (const py::builtin_exception&) {
        throw;
    } catch (const RenderCancelled& e) {
        raise_py(g_cancelled, e.what());
    } catch (const SessionClosed& e) {
        raise_py(PyExc_RuntimeError, e.what());
    } catch (const QueueFull& e) {
        raise_py(PyExc_RuntimeError, e.what());
    } catch (const std::bad_alloc&) {
        throw;
    } catch (const std::exception& e) {
        throw py::value_error(e.what());
    }
}

// ---- option readers (RawEnginePython.cpp:146-265) ----

bool has(const py::dict& d, const char* k) { return d.contains(k); }

void read_uint(const py::dict& d, const char* k, std::uint32_t& v) {
    if (!has(d, k)) return;
    const auto n = py::cast<unsigned long long>(d[k]);
    if (n > std::numeric_limits<std::uint32_t>::max()) raise_py(PyExc_OverflowError, "option exceeds uint32 range");
    v = std::uint32_t(n);
}

void read_float(const py::dict& d, const char* k, float& v) {
    if (has(d, k)) v = float(py::cast<double>(d[k]));
}

void read_levels(const py::dict& d, const char* k, std::array<std::uint16_t, 4>& levels) {
    if (!has(d, k)) return;
    const auto seq = py::cast<py::sequence>(d[k]);
    if (py::len(seq) != 4) throw py::value_error("site levels must contain exactly four values");
    for (size_t s = 0; s < 4; ++s) {
        const auto n = py::cast<unsigned long long>(seq[s]);
        if (n > 65535) throw py::value_error("site levels must be uint16 values");
        levels[s] = std::uint16_t(n);
    }
}

void read_color_transform(const py::dict& d, GraphRecipe& recipe) {
    if (!has(d, "camera_to_xyz_d50")) {
        if (has(d, "working_space")) throw py::value_error("working_space requires camera_to_xyz_d50");
        return;
    }
    const auto seq = py::cast<py::sequence>(d["camera_to_xyz_d50"]);
    if (py::len(seq) != 9) throw py::value_error("camera_to_xyz_d50 must contain nine numbers");
    CameraColorTransform t;
    for (size_t i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = py::cast<double>(seq[i]);
    if (has(d, "working_space")) {
        const auto name = py::cast<std::string>(d["working_space"]);
        if (name == "prophoto-d50") t.target = WorkingSpace::LinearProPhotoD50;
        else if (name == "rec2020-d65") t.target = WorkingSpace::LinearRec2020D65;
        else throw py::value_error("working_space must be prophoto-d50 or rec2020-d65");
    }
    recipe.camera_color = t;
}

void read_output_mode(const py::dict& d, GraphRecipe& recipe) {
    if (!has(d, "output_mode")) return;
    const auto name = py::cast<std::string>(d["output_mode"]);
    if (name == "legacy") recipe.output_mode = OutputMode::LegacyBounded;
    else if (name == "srgb-preview") recipe.output_mode = OutputMode::SrgbPreview;
    else if (name == "icc-display") recipe.output_mode = OutputMode::IccDisplay;
    else throw py::value_error("output_mode must be legacy, srgb-preview or icc-display");
}

void read_level(const py::dict& d, RenderLevel& level) {
    read_uint(d, "mip", level.mip);
    if (level.mip > 2) throw py::value_error("mip must be 0, 1 or 2");
    if (has(d, "quality")) {
        const auto name = py::cast<std::string>(d["quality"]);
        if (name == "preview") level.quality = RenderQuality::Preview;
        else if (name == "final") level.quality = RenderQuality::Final;
        else throw py::value_error("quality must be preview or final");
    }
}

py::dict options_of(const py::object& o) {
    if (o.is_none()) return py::dict();
    if (!py::isinstance<py::dict>(o)) throw py::type_error("options must be a dict");
    return py::reinterpret_borrow<py::dict>(o);
}

// Contiguous host bytes of a buffer-protocol object.
std::vector<std::uint8_t> copy_buffer(const py::object& src, size_t expected, const char* what) {
    py::buffer_info info = py::cast<py::buffer>(src).request();
    const size_t bytes = size_t(info.size) * size_t(info.itemsize);
    bool contiguous = true;
    ssize_t step = info.itemsize;
    for (ssize_t i = info.ndim - 1; i >= 0; --i) {
        if (info.shape[i] > 1 && info.strides[i] != step) contiguous = false;
        step *= info.shape[i];
    }
    if (!contiguous || bytes != expected) throw py::value_error(what);
    std::vector<std::uint8_t> out(bytes);
    std::memcpy(out.data(), info.ptr, bytes);
    return out;
}

py::tuple tile_result(const Tile& t) {
    if (t.rgb.size() > size_t(PY_SSIZE_T_MAX) / sizeof(float))
        throw std::length_error("requested ROI exceeds Python bytes capacity");
    return py::make_tuple(t.bounds.width, t.bounds.height,
                          py::bytes(reinterpret_cast<const char*>(t.rgb.data()), t.rgb.size() * sizeof(float)));
}

// ---- DLPack (the data-exchange ABI of dlpack.h, version 0.8) ----

struct DLDevice { std::int32_t device_type, device_id; };
struct DLDataType { std::uint8_t code, bits; std::uint16_t lanes; };
struct DLTensor {
    void* data;
    DLDevice device;
    std::int32_t ndim;
    DLDataType dtype;
    std::int64_t* shape;
    std::int64_t* strides;
    std::uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void* manager_ctx;
    void (*deleter)(DLManagedTensor*);
};
constexpr std::int32_t kDLROCM = 10;
constexpr std::uint8_t kDLFloat = 2;

struct DeviceResult {
    DLManagedTensor managed{};
    std::int64_t shape[3]{};
    void* memory = nullptr;
};

void free_result(DLManagedTensor* m) {
    auto* r = static_cast<DeviceResult*>(m->manager_ctx);
    if (r->memory) (void)hipFree(r->memory);
    delete r;
}

// Renders `output` into fresh device memory and wraps it as a "dltensor" capsule.
py::tuple render_capsule(const Node& output, Rect source_bounds, const RenderRequest& request) {
    auto r = std::make_unique<DeviceResult>();
    const Rect v = request.viewport;
    const size_t bytes = size_t(v.width) * v.height * 3 * sizeof(float);
    int device = 0;
    {
        py::gil_scoped_release unlocked;
        if (hipGetDevice(&device) != hipSuccess) throw std::runtime_error("no HIP device");
        const hipStream_t stream = detail::thread_stream();
        // From the stream-ordered pool; the capsule's deleter returns it.
        if (bytes && hipMallocAsync(&r->memory, bytes, stream) != hipSuccess) throw std::bad_alloc();
        try {
            detail::render_to_device(output, source_bounds, request, static_cast<float*>(r->memory), stream);
        } catch (...) {
            if (r->memory) (void)hipFreeAsync(r->memory, stream);
            r->memory = nullptr;
            throw;
        }
    }
    r->shape[0] = v.height;
    r->shape[1] = v.width;
    r->shape[2] = 3;
    r->managed.dl_tensor = {r->memory, {kDLROCM, device}, 3, {kDLFloat, 32, 1}, r->shape, nullptr, 0};
    r->managed.manager_ctx = r.get();
    r->managed.deleter = free_result;
    DLManagedTensor* m = &r.release()->managed;
    py::capsule cap(m, "dltensor", [](PyObject* o) {
        if (PyCapsule_IsValid(o, "dltensor")) {  // never consumed
            auto* t = static_cast<DLManagedTensor*>(PyCapsule_GetPointer(o, "dltensor"));
            t->deleter(t);
        }
    });
    return py::make_tuple(v.width, v.height, cap);
}

// Fault in the pages of a fresh host buffer in one call
// (MADV_POPULATE_WRITE, Linux 5.14+; elsewhere the copy faults them in),
// instead of one page fault per 4 KiB page during the copy.
void populate(char* p, size_t bytes) {
#ifdef MADV_POPULATE_WRITE
    const uintptr_t page = 4096, a = (reinterpret_cast<uintptr_t>(p) + page - 1) & ~(page - 1);
    const uintptr_t e = (reinterpret_cast<uintptr_t>(p) + bytes) & ~(page - 1);
    if (e > a) (void)madvise(reinterpret_cast<void*>(a), e - a, MADV_POPULATE_WRITE);
#else
    (void)p;
    (void)bytes;
#endif
}

// Host-bytes result of a render: compiled chains render on the device and
// are copied from device memory into the Python bytes object through the
// pinned staging buffers (no zero-filled intermediate Tile); other chains take
// Renderer::render_image.  The bytes object is not visible to Python until
// it is returned, so it is filled without the GIL.
py::tuple render_bytes(const Node& output, Rect source_bounds, const RenderRequest& request) {
    if (!detail::compilable(output)) {
        Tile t;
        {
            py::gil_scoped_release unlocked;
            t = Renderer{}.render_image(output, source_bounds, request);
        }
        return tile_result(t);
    }
    const Rect v = request.viewport;
    const size_t bytes = size_t(v.width) * v.height * 3 * sizeof(float);
    if (bytes > size_t(PY_SSIZE_T_MAX)) throw std::length_error("requested ROI exceeds Python bytes capacity");
    PyObject* obj = PyBytes_FromStringAndSize(nullptr, Py_ssize_t(bytes));
    if (!obj) throw py::error_already_set();
    py::bytes result = py::reinterpret_steal<py::bytes>(obj);
    char* host = PyBytes_AS_STRING(obj);
    {
        py::gil_scoped_release unlocked;
        const hipStream_t stream = detail::thread_stream();
        void* dev = nullptr;
        if (bytes && hipMallocAsync(&dev, bytes, stream) != hipSuccess) throw std::bad_alloc();
        try {
            detail::render_to_device(output, source_bounds, request, static_cast<float*>(dev), stream);
            populate(host, bytes);
            detail::copy_to_host(host, dev, bytes, stream);
        } catch (...) {
            if (dev) (void)hipFreeAsync(dev, stream);
            throw;
        }
        if (dev) detail::hip(hipFreeAsync(dev, stream), "hipFreeAsync");
    }
    return py::make_tuple(v.width, v.height, result);
}

// ---- one-shot RAW render (RawEnginePython.cpp:387-519) ----

struct OneShot {
    std::optional<RawImage> image;
    std::optional<ImageGraph> graph;
    RenderRequest request;
};

OneShot prepare_render(const py::object& source, unsigned width, unsigned height, const py::object& options_obj) {
    const py::dict o = options_of(options_obj);
    std::uint32_t black = 0, white = 65535, pattern = 0, tile = 256, stride = 0, px = 0, py_ = 0;
    Rect area;
    GraphRecipe recipe;
    RenderLevel level;
    const bool explicit_roi = has(o, "x") || has(o, "y") || has(o, "roi_width") || has(o, "roi_height");
    read_uint(o, "black_level", black);
    read_uint(o, "white_level", white);
    read_uint(o, "pattern", pattern);
    read_uint(o, "row_stride_samples", stride);
    read_uint(o, "cfa_phase_x", px);
    read_uint(o, "cfa_phase_y", py_);
    read_uint(o, "active_x", area.x);
    read_uint(o, "active_y", area.y);
    read_uint(o, "active_width", area.width);
    read_uint(o, "active_height", area.height);
    read_uint(o, "tile_size", tile);
    read_float(o, "red_gain", recipe.red_gain);
    read_float(o, "green_gain", recipe.green_gain);
    read_float(o, "blue_gain", recipe.blue_gain);
    read_float(o, "exposure_stops", recipe.exposure_stops);
    read_float(o, "tone_shoulder", recipe.tone_shoulder);
    read_float(o, "tone_gamma", recipe.tone_gamma);
    read_color_transform(o, recipe);
    read_output_mode(o, recipe);
    read_level(o, level);
    if (has(o, "input_profile")) throw std::invalid_argument("Bayer input does not accept input_profile");
    if (has(o, "output_profile") && !o["output_profile"].is_none())
        throw std::invalid_argument("ICC profiles require a build with a colour-management engine");
    if (recipe.output_mode == OutputMode::IccDisplay)
        throw std::invalid_argument("icc-display output requires output_profile");
    if (!(level.mip == 0 && level.quality == RenderQuality::Final) &&
        !(level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview && recipe.camera_color &&
          recipe.output_mode == OutputMode::SrgbPreview))
        throw std::invalid_argument("RAW preview requires mip 1/2, preview quality, camera calibration and sRGB output");
    if (has(o, "crop")) throw std::invalid_argument("RAW render does not accept raster crop; use active area or ROI");
    if (has(o, "resize") || has(o, "resize_filter")) throw std::invalid_argument("RAW render does not accept raster resize");
    if (has(o, "rotate") || has(o, "flip_horizontal") || has(o, "flip_vertical"))
        throw std::invalid_argument("RAW render does not accept raster orientation options");
    if (black > 65535 || white > 65535 || pattern > 3 || px > 1 || py_ > 1)
        throw std::invalid_argument("invalid levels, Bayer pattern, or CFA phase");
    RawMetadata m;
    m.width = width;
    m.height = height;
    m.row_stride_samples = stride;
    m.pattern = BayerPattern(pattern);
    m.cfa_phase_x = std::uint8_t(px);
    m.cfa_phase_y = std::uint8_t(py_);
    m.active_area = area;
    m.black_levels.fill(std::uint16_t(black));
    m.white_levels.fill(std::uint16_t(white));
    read_levels(o, "black_levels", m.black_levels);
    read_levels(o, "white_levels", m.white_levels);
    Rect roi{0, 0, width, height};
    if (!explicit_roi && area.width && area.height) roi = area;
    if (level.mip) {
        const uint32_t s = 1u << level.mip;
        const uint32_t aw = area.width ? area.width : width, ah = area.height ? area.height : height;
        roi = {0, 0, aw / s + (aw % s != 0), ah / s + (ah % s != 0)};
    }
    read_uint(o, "x", roi.x);
    read_uint(o, "y", roi.y);
    read_uint(o, "roi_width", roi.width);
    read_uint(o, "roi_height", roi.height);
    const uint64_t count = uint64_t(stride ? stride : width) * height;
    if (!width || !height || count > uint64_t(PY_SSIZE_T_MAX) / 2)
        throw std::invalid_argument("bayer must contain row_stride_samples*height native-endian uint16 samples");
    const auto raw = copy_buffer(source, size_t(count) * 2,
                                 "bayer must contain row_stride_samples*height native-endian uint16 samples");
    std::vector<std::uint16_t> samples(count);
    std::memcpy(samples.data(), raw.data(), raw.size());
    OneShot s;
    {
        py::gil_scoped_release unlocked;
        s.image.emplace(m, std::move(samples));
        s.graph.emplace(*s.image, recipe);
    }
    s.request = RenderRequest{roi, tile, level};
    return s;
}

py::tuple render(const py::object& bayer, unsigned width, unsigned height, const py::object& options) {
    return guarded([&] {
        OneShot s = prepare_render(bayer, width, height, options);
        return render_bytes(s.graph->output(), s.graph->source_bounds(), s.request);
    });
}

py::tuple render_dlpack(const py::object& bayer, unsigned width, unsigned height, const py::object& options) {
    return guarded([&] {
        OneShot s = prepare_render(bayer, width, height, options);
        return render_capsule(s.graph->output(), s.graph->source_bounds(), s.request);
    });
}

// ---- one-shot raster render (RawEnginePython.cpp:521-639) ----

py::tuple render_raster(const py::object& rgb, unsigned width, unsigned height, const py::object& options_obj) {
    return guarded([&] {
        const py::dict o = options_of(options_obj);
        std::uint32_t stride = 0;
        read_uint(o, "row_stride_pixels", stride);
        for (const char* k : {"crop", "resize", "resize_filter", "rotate", "flip_horizontal", "flip_vertical"})
            if (has(o, k) && !o[k].is_none())
                throw std::invalid_argument(std::string("raster option '") + k +
                                            "' is outside the device render path of this build");
        if (has(o, "input_profile") || (has(o, "output_profile") && !o["output_profile"].is_none()))
            throw std::invalid_argument("ICC profiles require a build with a colour-management engine");
        GraphRecipe recipe;
        RenderRequest request;
        read_level(o, request.level);
        const uint32_t s = 1u << request.level.mip;
        request.viewport = {0, 0, width / s + (width % s != 0), height / s + (height % s != 0)};
        read_uint(o, "x", request.viewport.x);
        read_uint(o, "y", request.viewport.y);
        read_uint(o, "roi_width", request.viewport.width);
        read_uint(o, "roi_height", request.viewport.height);
        read_uint(o, "tile_size", request.tile_size);
        read_float(o, "exposure_stops", recipe.exposure_stops);
        read_float(o, "tone_shoulder", recipe.tone_shoulder);
        read_float(o, "tone_gamma", recipe.tone_gamma);
        read_output_mode(o, recipe);
        if (recipe.output_mode == OutputMode::IccDisplay)
            throw std::invalid_argument("icc-display output requires output_profile");
        for (const char* k : {"red_gain", "green_gain", "blue_gain", "camera_to_xyz_d50"})
            if (has(o, k)) throw std::invalid_argument("raster input does not accept RAW calibration controls");
        if (!has(o, "working_space")) throw py::value_error("raster working_space is required");
        const auto name = py::cast<std::string>(o["working_space"]);
        WorkingSpace space;
        if (name == "prophoto-d50") space = WorkingSpace::LinearProPhotoD50;
        else if (name == "rec2020-d65") space = WorkingSpace::LinearRec2020D65;
        else throw py::value_error("working_space must be prophoto-d50 or rec2020-d65");
        const uint64_t rows = uint64_t(stride ? stride : width) * height;
        if (!width || !height || rows > uint64_t(PY_SSIZE_T_MAX) / 12)
            throw std::invalid_argument("rgb must contain exactly row_stride_pixels*height RGB pixels");
        const auto raw = copy_buffer(rgb, size_t(rows) * 12,
                                     "rgb must contain exactly row_stride_pixels*height RGB pixels in the declared "
                                     "source format");
        std::vector<float> px(rows * 3);
        std::memcpy(px.data(), raw.data(), raw.size());
        std::optional<ImageGraph> graph;
        {
            py::gil_scoped_release unlocked;
            graph.emplace(RasterImage({width, height, stride, space}, std::move(px)), recipe);
        }
        return render_bytes(graph->output(), graph->source_bounds(), request);
    });
}

// ---- sessions and jobs (RawEnginePython.cpp:735-1001, 2018-2256) ----

struct Job {
    std::shared_future<Tile> result;
    std::shared_ptr<CancellationToken> cancellation = std::make_shared<CancellationToken>();
    std::uint64_t total_tiles = 0;
    std::shared_ptr<TileScheduler> scheduler;  // keeps the workers alive
    ~Job() { cancellation->cancel(); }
};

class RenderJob {
public:
    explicit RenderJob(std::shared_ptr<Job> j) : j_(std::move(j)) {}
    ~RenderJob() {
        py::gil_scoped_release unlocked;
        j_.reset();
    }
    std::string output_kind() const { return "rgb"; }
    bool done() const { return j_->result.wait_for(std::chrono::seconds(0)) == std::future_status::ready; }
    void cancel() { j_->cancellation->cancel(); }
    // Tiles are produced by whole-viewport device launches: the count moves
    // from 0 to the total when the render has completed.
    py::dict progress() const {
        bool ok = false;
        if (done()) {
            try {
                (void)j_->result.get();
                ok = true;
            } catch (...) {
            }
        }
        py::dict d;
        d["completed_tiles"] = ok ? j_->total_tiles : 0;
        d["total_tiles"] = j_->total_tiles;
        return d;
    }
    py::tuple result(const py::object& timeout) {
        double seconds = 0;
        if (!timeout.is_none()) {
            seconds = py::cast<double>(timeout);
            if (!std::isfinite(seconds) || seconds < 0)
                throw py::value_error("timeout must be finite and nonnegative, or None");
        }
        bool ready = true;
        {
            py::gil_scoped_release unlocked;
            if (timeout.is_none()) j_->result.wait();
            else ready = j_->result.wait_for(std::chrono::duration<double>(seconds)) == std::future_status::ready;
        }
        if (!ready) raise_py(PyExc_TimeoutError, "render job did not finish before the timeout");
        try {
            return tile_result(j_->result.get());
        } catch (const RenderCancelled& e) {
            raise_py(g_cancelled, e.what());
        } catch (const std::invalid_argument& e) {
            throw py::value_error(e.what());
        } catch (const std::domain_error& e) {
            throw py::value_error(e.what());
        }
    }

private:
    std::shared_ptr<Job> j_;
};

constexpr const char* kSourceId = "00000000-0000-0000-0000-000000000001";

class Session {
public:
    Session(RawImage image, size_t budget, size_t workers, size_t pending, RawDemosaicIdentity demosaic)
        : raw_metadata_(image.metadata()), bounds_(image.metadata().active_area),
          cache_(std::make_shared<TileCache>(budget)), cache_bytes_(budget), workers_(workers), max_pending_(pending) {
        if (!workers || workers > 64 || !pending)
            throw std::invalid_argument("session needs 1-64 workers and a positive pending budget");
        source_.id = kSourceId;
        source_.kind = EditSourceKind::DecodedBayerU16;
        source_.content_sha256 = image.fingerprint();
        source_.demosaic = demosaic;
        node_ = std::make_shared<RawUnpackNode>(std::move(image), std::move(demosaic));
    }
    Session(RasterImage image, size_t budget, size_t workers, size_t pending)
        : bounds_{0, 0, image.width(), image.height()}, cache_(std::make_shared<TileCache>(budget)),
          cache_bytes_(budget), workers_(workers), max_pending_(pending) {
        if (!workers || workers > 64 || !pending)
            throw std::invalid_argument("session needs 1-64 workers and a positive pending budget");
        source_.id = kSourceId;
        source_.kind = EditSourceKind::SceneLinearRasterF32;
        source_.working_space = image.metadata().working_space;
        source_.content_sha256 = fingerprint_raster_source(image);
        node_ = std::make_shared<RasterSourceNode>(std::move(image));
    }
    ~Session() { close(); }

    // The recipe as a saved graph (RawEnginePython.cpp:832-931).
    ExecutableEditGraph recipe_graph(const GraphRecipe& recipe) const {
        if (closed_.load()) throw SessionClosed();
        EditManifest m;
        if (raw_metadata_) m.format_version = 3;
        m.working_space = raw_metadata_
                              ? (recipe.camera_color ? recipe.camera_color->target : WorkingSpace::LinearProPhotoD50)
                              : *source_.working_space;
        m.sources.push_back(source_);
        const auto working = m.working_space == WorkingSpace::LinearProPhotoD50 ? EditDomain::SceneLinearProPhotoD50
                                                                               : EditDomain::SceneLinearRec2020D65;
        auto domain = raw_metadata_ ? EditDomain::CameraLinear : working;
        std::string upstream = source_.id;
        auto append = [&](const char* id, const char* type, EditDomain out, EditValue::Object params = {}) {
            EditOperation op;
            op.id = id;
            op.type_id = type;
            op.processing_version = kCurrentEditProcessingVersion;
            op.input_domain = domain;
            op.output_domain = out;
            op.inputs.emplace("image", upstream);
            op.parameters = std::move(params);
            upstream = op.id;
            domain = out;
            m.operations.push_back(std::move(op));
        };
        if (raw_metadata_) {
            append("00000000-0000-0000-0000-000000000010", "rawengine.white_balance", domain,
                   {{"red_gain", EditValue{double(recipe.red_gain)}},
                    {"green_gain", EditValue{double(recipe.green_gain)}},
                    {"blue_gain", EditValue{double(recipe.blue_gain)}}});
            append("00000000-0000-0000-0000-000000000002", "rawengine.exposure", domain,
                   {{"stops", EditValue{double(recipe.exposure_stops)}}});
            if (recipe.camera_color) {
                EditValue::Array mat;
                for (double v : recipe.camera_color->camera_to_xyz_d50) mat.push_back(EditValue{v});
                append("00000000-0000-0000-0000-000000000011", "rawengine.camera_to_working", working,
                       {{"matrix", EditValue{std::move(mat)}}});
            }
        } else {
            append("00000000-0000-0000-0000-000000000002", "rawengine.exposure", domain,
                   {{"stops", EditValue{double(recipe.exposure_stops)}}});
        }
        if (recipe.output_mode == OutputMode::IccDisplay)
            throw std::invalid_argument("icc-display output requires the session output_profile");
        const bool display = recipe.output_mode != OutputMode::LegacyBounded;
        if (display) append("00000000-0000-0000-0000-000000000003", "rawengine.working_to_srgb", EditDomain::SceneLinearSrgb);
        append("00000000-0000-0000-0000-000000000004", "rawengine.tone_curve",
               display ? EditDomain::DisplayLinearSrgb : EditDomain::ToneMappedUnmanaged,
               {{"shoulder", EditValue{double(recipe.tone_shoulder)}}, {"gamma", EditValue{double(recipe.tone_gamma)}}});
        if (recipe.output_mode == OutputMode::SrgbPreview)
            append("00000000-0000-0000-0000-000000000005", "rawengine.srgb_encode", EditDomain::DisplayEncodedSrgb);
        else
            append("00000000-0000-0000-0000-000000000006", "rawengine.output_clip", EditDomain::UnmanagedBounded);
        m.output_id = upstream;
        return ExecutableEditGraph(std::move(m), {{source_, node_, bounds_}}, nullptr, cache_);
    }

    ExecutableEditGraph manifest_graph(const std::string& json) const {
        if (closed_.load()) throw SessionClosed();
        return ExecutableEditGraph(parse_edit_manifest(json), {{source_, node_, bounds_}}, nullptr, cache_);
    }

    // Manifest requests: ROI, tile size and level only (RawEnginePython.cpp:1009-1036).
    static RenderRequest manifest_request(const py::dict& o, Rect bounds) {
        for (auto item : o) {
            if (!py::isinstance<py::str>(item.first)) throw py::type_error("manifest request keys must be strings");
            const auto k = py::cast<std::string>(item.first);
            if (k != "x" && k != "y" && k != "roi_width" && k != "roi_height" && k != "tile_size" && k != "mip" &&
                k != "quality")
                throw py::value_error(
                    "manifest options accept only ROI, tile_size, mip and quality; edits belong in the manifest");
        }
        RenderRequest r;
        read_level(o, r.level);
        const uint32_t s = 1u << r.level.mip;
        r.viewport = {r.level.mip ? 0 : bounds.x, r.level.mip ? 0 : bounds.y, bounds.width / s + (bounds.width % s != 0),
                      bounds.height / s + (bounds.height % s != 0)};
        read_uint(o, "x", r.viewport.x);
        read_uint(o, "y", r.viewport.y);
        read_uint(o, "roi_width", r.viewport.width);
        read_uint(o, "roi_height", r.viewport.height);
        read_uint(o, "tile_size", r.tile_size);
        return r;
    }

    // RawEnginePython.cpp:1038-1105.
    std::pair<std::optional<ExecutableEditGraph>, RenderRequest> prepare(const py::object& options,
                                                                         const py::object& manifest) const {
        const py::dict o = options_of(options);
        std::optional<ExecutableEditGraph> g;
        if (!manifest.is_none()) {
            const std::string json = py::cast<std::string>(manifest);
            {
                py::gil_scoped_release unlocked;
                g.emplace(manifest_graph(json));
            }
            return {std::move(g), manifest_request(o, g->output_bounds())};
        }
        RenderRequest r;
        GraphRecipe recipe;
        if (raw_metadata_) {
            static const char* allowed[] = {"x", "y", "roi_width", "roi_height", "tile_size", "mip", "quality",
                                            "red_gain", "green_gain", "blue_gain", "camera_to_xyz_d50",
                                            "working_space", "exposure_stops", "tone_shoulder", "tone_gamma",
                                            "output_mode"};
            for (auto item : o) {
                const auto k = py::cast<std::string>(item.first);
                if (std::find_if(std::begin(allowed), std::end(allowed), [&](const char* a) { return k == a; }) ==
                    std::end(allowed))
                    throw std::invalid_argument(
                        "unknown RAW recipe option; sensor metadata and session budgets are fixed");
            }
            read_float(o, "red_gain", recipe.red_gain);
            read_float(o, "green_gain", recipe.green_gain);
            read_float(o, "blue_gain", recipe.blue_gain);
            read_color_transform(o, recipe);
        } else {
            for (const char* k : {"working_space", "row_stride_pixels", "red_gain", "green_gain", "blue_gain",
                                  "camera_to_xyz_d50", "cache_bytes", "workers", "max_pending", "input_profile",
                                  "output_profile"})
                if (has(o, k))
                    throw std::invalid_argument("session render options cannot change source, RAW calibration or budgets");
            for (const char* k : {"crop", "resize", "resize_filter", "rotate", "flip_horizontal", "flip_vertical"})
                if (has(o, k) && !o[k].is_none())
                    throw std::invalid_argument(std::string("raster option '") + k +
                                                "' is outside the device render path of this build");
        }
        read_float(o, "exposure_stops", recipe.exposure_stops);
        read_float(o, "tone_shoulder", recipe.tone_shoulder);
        read_float(o, "tone_gamma", recipe.tone_gamma);
        read_output_mode(o, recipe);
        read_level(o, r.level);
        const uint32_t s = 1u << r.level.mip;
        if (raw_metadata_)
            r.viewport = {r.level.mip ? 0 : bounds_.x, r.level.mip ? 0 : bounds_.y,
                          bounds_.width / s + (bounds_.width % s != 0), bounds_.height / s + (bounds_.height % s != 0)};
        else
            r.viewport = {0, 0, bounds_.width / s + (bounds_.width % s != 0), bounds_.height / s + (bounds_.height % s != 0)};
        read_uint(o, "x", r.viewport.x);
        read_uint(o, "y", r.viewport.y);
        read_uint(o, "roi_width", r.viewport.width);
        read_uint(o, "roi_height", r.viewport.height);
        read_uint(o, "tile_size", r.tile_size);
        {
            py::gil_scoped_release unlocked;
            g.emplace(recipe_graph(recipe));
        }
        return {std::move(g), r};
    }

    py::tuple render(const py::object& options, const py::object& manifest) const {
        return guarded([&] {
            auto [g, r] = prepare(options, manifest);
            return render_bytes(g->output(), g->output_bounds(), r);
        });
    }

    py::tuple render_dlpack(const py::object& options, const py::object& manifest) const {
        return guarded([&] {
            auto [g, r] = prepare(options, manifest);
            return render_capsule(g->output(), g->output_bounds(), r);
        });
    }

    std::string export_manifest(const py::object& options) const {
        return guarded([&] {
            auto [g, r] = prepare(options, py::none());
            (void)r;
            py::gil_scoped_release unlocked;
            return serialize_edit_manifest(g->manifest());
        });
    }

    py::dict required_regions(const py::object& manifest, const py::object& options) const {
        return guarded([&] {
            auto [g, r] = prepare(options, manifest);
            if (!r.tile_size) throw std::invalid_argument("tile size must be positive");
            std::map<std::string, Rect> regions;
            {
                py::gil_scoped_release unlocked;
                regions = g->required_source_regions(r.viewport, r.level);
            }
            py::dict d;
            for (const auto& [id, rc] : regions) d[py::str(id)] = py::make_tuple(rc.x, rc.y, rc.width, rc.height);
            return d;
        });
    }

    py::dict histogram(const py::object& manifest, const py::object& options, unsigned bins, double lower,
                       double upper) const {
        return guarded([&] {
            if (bins == 0 || bins > 65536) throw py::value_error("bins must be in 1..65536");
            auto [g, r] = prepare(options, manifest);
            RgbHistogram h;
            {
                py::gil_scoped_release unlocked;
                h = histogram_rgb(*g, r, HistogramOptions{bins, lower, upper});
            }
            py::list channels;
            for (const auto& c : h.channels) {
                py::dict ch;
                ch["counts"] = py::cast(c.counts);
                ch["underflow"] = c.underflow;
                ch["overflow"] = c.overflow;
                ch["nonfinite"] = c.nonfinite;
                ch["minimum"] = c.minimum ? py::cast(double(*c.minimum)) : py::none();
                ch["maximum"] = c.maximum ? py::cast(double(*c.maximum)) : py::none();
                channels.append(ch);
            }
            py::dict d;
            d["version"] = 1;
            d["bounds"] = py::make_tuple(r.viewport.x, r.viewport.y, r.viewport.width, r.viewport.height);
            d["mip"] = r.level.mip;
            d["quality"] = r.level.quality == RenderQuality::Preview ? "preview" : "final";
            d["bins"] = bins;
            d["lower"] = lower;
            d["upper"] = upper;
            d["pixel_count"] = h.pixel_count;
            d["channels"] = channels;
            return d;
        });
    }

    // RawEnginePython.cpp:1792-1854: one callback per statistics tile with
    // the moments as bytes (float64 mean / variance, uint32 finite counts).
    py::dict analyze_local(const py::object& manifest, const py::object& callback, const py::object& options,
                           const py::object& radius_obj) const {
        return guarded([&] {
            if (!PyCallable_Check(callback.ptr())) throw py::type_error("callback must be callable");
            uint32_t radius = 3;
            if (!radius_obj.is_none()) {
                if (!py::isinstance<py::int_>(radius_obj) || py::isinstance<py::bool_>(radius_obj))
                    throw py::type_error("radius must be an integer");
                const auto n = py::cast<long long>(radius_obj);
                if (n < 0 || n > 8) throw py::value_error("radius must be in 0..8");
                radius = uint32_t(n);
            }
            auto [g, r] = prepare(options, manifest);
            uint64_t tiles = 0, pixels = 0;
            {
                py::gil_scoped_release unlocked;
                analyze_local_rgb(*g, r, [&](const LocalStatisticsTile& t) {
                    py::gil_scoped_acquire locked;
                    py::dict payload;
                    payload["version"] = 1;
                    payload["bounds"] = py::make_tuple(t.bounds.x, t.bounds.y, t.bounds.width, t.bounds.height);
                    payload["radius"] = radius;
                    payload["mip"] = r.level.mip;
                    payload["quality"] = r.level.quality == RenderQuality::Preview ? "preview" : "final";
                    payload["mean_f64"] = py::bytes(reinterpret_cast<const char*>(t.mean.data()),
                                                    t.mean.size() * sizeof(double));
                    payload["variance_f64"] = py::bytes(reinterpret_cast<const char*>(t.variance.data()),
                                                        t.variance.size() * sizeof(double));
                    payload["finite_count_u32"] = py::bytes(reinterpret_cast<const char*>(t.finite_count.data()),
                                                            t.finite_count.size() * sizeof(uint32_t));
                    callback(payload);
                    ++tiles;
                    pixels += uint64_t(t.bounds.width) * t.bounds.height;
                }, radius);
            }
            py::dict d;
            d["completed_tiles"] = tiles;
            d["pixel_count"] = pixels;
            return d;
        });
    }

    RenderJob submit(const py::object& options, const py::object& manifest, const std::string& priority,
                     const py::object& group_obj) {
        return guarded([&] {
            RenderPriority p;
            if (priority == "background") p = RenderPriority::Background;
            else if (priority == "normal") p = RenderPriority::Normal;
            else if (priority == "interactive") p = RenderPriority::Interactive;
            else throw py::value_error("priority must be background, normal or interactive");
            std::string group;
            if (!group_obj.is_none()) {
                group = py::cast<std::string>(group_obj);
                if (group.empty()) throw std::invalid_argument("latest request group must not be empty");
            }
            auto [g, r] = prepare(options, manifest);
            if (!r.tile_size) throw std::invalid_argument("tile size must be positive");
            auto job = std::make_shared<Job>();
            const uint64_t cols = r.viewport.width / r.tile_size + (r.viewport.width % r.tile_size != 0);
            const uint64_t rows = r.viewport.height / r.tile_size + (r.viewport.height % r.tile_size != 0);
            job->total_tiles = cols * rows;
            {
                py::gil_scoped_release unlocked;
                std::lock_guard lock(scheduler_mutex_);
                if (closed_.load()) throw SessionClosed();
                if (!scheduler_) scheduler_ = std::make_shared<TileScheduler>(workers_, max_pending_);
                std::erase_if(jobs_, [](const auto& w) {
                    auto t = w.lock();
                    return !t || t->is_cancelled();
                });
                jobs_.push_back(job->cancellation);
                job->scheduler = scheduler_;
                try {
                    job->result = (group.empty() ? scheduler_->submit(g->output_handle(), g->output_bounds(), r, p,
                                                                      job->cancellation)
                                                 : scheduler_->submit_latest(group, g->output_handle(),
                                                                             g->output_bounds(), r, p,
                                                                             job->cancellation))
                                      .share();
                } catch (const std::length_error& e) {
                    throw QueueFull(e.what());
                }
            }
            return RenderJob(job);
        });
    }

    py::dict source_info() const {
        static const char* hexdig = "0123456789abcdef";
        std::string digest;
        for (auto b : source_.content_sha256) { digest += hexdig[b >> 4]; digest += hexdig[b & 15]; }
        py::dict d;
        d["id"] = source_.id;
        d["content_sha256"] = digest;
        if (raw_metadata_) {
            const auto& m = *raw_metadata_;
            d["kind"] = "decoded_bayer_u16";
            py::dict dm;
            dm["algorithm"] = source_.demosaic->algorithm;
            dm["processing_version"] = source_.demosaic->processing_version;
            d["demosaic"] = dm;
            d["width"] = m.width;
            d["height"] = m.height;
            d["row_stride_samples"] = m.row_stride_samples;
            d["pattern"] = unsigned(m.pattern);
            d["cfa_phase_x"] = unsigned(m.cfa_phase_x);
            d["cfa_phase_y"] = unsigned(m.cfa_phase_y);
            d["active_area"] = py::make_tuple(m.active_area.x, m.active_area.y, m.active_area.width, m.active_area.height);
            d["black_levels"] = py::make_tuple(m.black_levels[0], m.black_levels[1], m.black_levels[2], m.black_levels[3]);
            d["white_levels"] = py::make_tuple(m.white_levels[0], m.white_levels[1], m.white_levels[2], m.white_levels[3]);
        } else {
            d["kind"] = "scene_linear_raster_f32";
            d["working_space"] = *source_.working_space == WorkingSpace::LinearProPhotoD50 ? "linear_prophoto_d50"
                                                                                          : "linear_rec2020_d65";
            d["width"] = bounds_.width;
            d["height"] = bounds_.height;
        }
        return d;
    }

    py::dict cache_stats() const {
        const auto s = cache_->stats();
        py::dict d;
        d["entries"] = s.entries;
        d["used_bytes"] = s.used_bytes;
        d["hits"] = s.hits;
        d["misses"] = s.misses;
        d["budget_bytes"] = cache_bytes_;
        return d;
    }

    void clear_cache() { cache_->clear(); }

    // Rejects new work and cancels unfinished jobs; idempotent.
    void close() {
        std::shared_ptr<TileScheduler> owned;
        {
            std::lock_guard lock(scheduler_mutex_);
            closed_.store(true);
            for (const auto& w : jobs_)
                if (auto t = w.lock()) t->cancel();
            jobs_.clear();
            owned = std::move(scheduler_);
        }
    }

private:
    EditSource source_;
    std::shared_ptr<const Node> node_;
    std::optional<RawMetadata> raw_metadata_;
    Rect bounds_{};
    std::shared_ptr<TileCache> cache_;
    size_t cache_bytes_;
    size_t workers_, max_pending_;
    std::mutex scheduler_mutex_;
    std::shared_ptr<TileScheduler> scheduler_;
    std::vector<std::weak_ptr<CancellationToken>> jobs_;
    std::atomic<bool> closed_{false};
};

struct RawSession : Session {
    using Session::Session;
};
struct RasterSession : Session {
    using Session::Session;
};

RawDemosaicIdentity demosaic_of(const py::object& d) {
    RawDemosaicIdentity id;
    if (d.is_none()) return id;
    if (!py::isinstance<py::dict>(d)) throw py::type_error("demosaic must be a dict or None");
    const py::dict dd = py::reinterpret_borrow<py::dict>(d);
    if (py::len(dd) != 2 || !has(dd, "algorithm") || !has(dd, "processing_version"))
        throw std::invalid_argument("demosaic requires exactly algorithm and processing_version");
    if (!py::isinstance<py::str>(dd["algorithm"]) || !py::isinstance<py::int_>(dd["processing_version"]) ||
        py::isinstance<py::bool_>(dd["processing_version"]))
        throw py::type_error("demosaic requires a string algorithm and integer processing_version");
    id.algorithm = py::cast<std::string>(dd["algorithm"]);
    const auto v = py::cast<unsigned long long>(dd["processing_version"]);
    if (v > std::numeric_limits<std::uint32_t>::max()) throw std::invalid_argument("invalid demosaic processing_version");
    id.processing_version = std::uint32_t(v);
    validate_raw_demosaic(id);
    return id;
}

// RawEnginePython.cpp:1302-1402.
std::unique_ptr<RawSession> raw_session(const py::object& bayer, unsigned width, unsigned height,
                                     const py::object& metadata, size_t cache_bytes, size_t workers,
                                     size_t max_pending, const py::object& demosaic, const py::object& output_profile) {
    return guarded([&] {
        if (!output_profile.is_none())
            throw std::invalid_argument("ICC profiles require a build with a colour-management engine");
        const auto id = demosaic_of(demosaic);
        RawMetadata m;
        m.width = width;
        m.height = height;
        if (!metadata.is_none()) {
            if (!py::isinstance<py::dict>(metadata)) throw py::type_error("metadata must be a dict or None");
            const py::dict o = py::reinterpret_borrow<py::dict>(metadata);
            static const char* allowed[] = {"row_stride_samples", "pattern", "cfa_phase_x", "cfa_phase_y",
                                            "active_x", "active_y", "active_width", "active_height",
                                            "black_level", "white_level", "black_levels", "white_levels"};
            for (auto item : o) {
                const auto k = py::cast<std::string>(item.first);
                if (std::find_if(std::begin(allowed), std::end(allowed), [&](const char* a) { return k == a; }) ==
                    std::end(allowed))
                    throw std::invalid_argument("unknown RAW metadata field; editing controls belong in render options");
                if (k != "black_levels" && k != "white_levels" &&
                    (!py::isinstance<py::int_>(item.second) || py::isinstance<py::bool_>(item.second)))
                    throw py::type_error("RAW metadata dimensions, phase, pattern and levels must be integers");
            }
            std::uint32_t black = 0, white = 65535, pattern = 0, px = 0, py_ = 0;
            read_uint(o, "row_stride_samples", m.row_stride_samples);
            read_uint(o, "pattern", pattern);
            read_uint(o, "cfa_phase_x", px);
            read_uint(o, "cfa_phase_y", py_);
            read_uint(o, "active_x", m.active_area.x);
            read_uint(o, "active_y", m.active_area.y);
            read_uint(o, "active_width", m.active_area.width);
            read_uint(o, "active_height", m.active_area.height);
            read_uint(o, "black_level", black);
            read_uint(o, "white_level", white);
            if (black > 65535 || white > 65535 || pattern > 3 || px > 1 || py_ > 1)
                throw std::invalid_argument("invalid RAW levels, pattern or CFA phase");
            m.pattern = BayerPattern(pattern);
            m.cfa_phase_x = std::uint8_t(px);
            m.cfa_phase_y = std::uint8_t(py_);
            m.black_levels.fill(std::uint16_t(black));
            m.white_levels.fill(std::uint16_t(white));
            read_levels(o, "black_levels", m.black_levels);
            read_levels(o, "white_levels", m.white_levels);
        }
        const uint32_t stride = m.row_stride_samples ? m.row_stride_samples : width;
        const uint64_t count = uint64_t(stride) * height;
        if (!width || !height || stride < width || count > uint64_t(PY_SSIZE_T_MAX) / 2)
            throw std::invalid_argument(
                "bayer must contain row_stride_samples*height native-endian uint16 samples with valid dimensions/stride");
        const auto raw = copy_buffer(bayer, size_t(count) * 2,
                                     "bayer must contain row_stride_samples*height native-endian uint16 samples with "
                                     "valid dimensions/stride");
        std::vector<std::uint16_t> samples(count);
        std::memcpy(samples.data(), raw.data(), raw.size());
        py::gil_scoped_release unlocked;
        return std::make_unique<RawSession>(RawImage(m, std::move(samples)), cache_bytes, workers, max_pending, id);
    });
}

std::unique_ptr<RasterSession> raster_session(const py::object& rgb, unsigned width, unsigned height,
                                        const std::string& working_space, unsigned row_stride_pixels,
                                        size_t cache_bytes, size_t workers, size_t max_pending) {
    return guarded([&] {
        WorkingSpace space;
        if (working_space == "prophoto-d50") space = WorkingSpace::LinearProPhotoD50;
        else if (working_space == "rec2020-d65") space = WorkingSpace::LinearRec2020D65;
        else throw py::value_error("working_space must be prophoto-d50 or rec2020-d65");
        const uint64_t rows = uint64_t(row_stride_pixels ? row_stride_pixels : width) * height;
        if (!width || !height || rows > uint64_t(PY_SSIZE_T_MAX) / 12)
            throw std::invalid_argument("rgb must contain exactly row_stride_pixels*height RGB pixels");
        const auto raw = copy_buffer(rgb, size_t(rows) * 12, "rgb must contain exactly row_stride_pixels*height RGB pixels");
        std::vector<float> px(rows * 3);
        std::memcpy(px.data(), raw.data(), raw.size());
        py::gil_scoped_release unlocked;
        return std::make_unique<RasterSession>(RasterImage({width, height, row_stride_pixels, space}, std::move(px)),
                                               cache_bytes, workers, max_pending);
    });
}

template <class T, class C> void bind_session(C cls) {
    cls.def("render", [](const T& s, const py::object& o) { return s.render(o, py::none()); },
            py::arg("options") = py::none())
        .def("render_dlpack", [](const T& s, const py::object& o) { return s.render_dlpack(o, py::none()); },
             py::arg("options") = py::none())
        .def("render_manifest", [](const T& s, const py::object& mf, const py::object& o) { return s.render(o, mf); },
             py::arg("manifest"), py::arg("options") = py::none())
        .def("render_manifest_dlpack",
             [](const T& s, const py::object& mf, const py::object& o) { return s.render_dlpack(o, mf); },
             py::arg("manifest"), py::arg("options") = py::none())
        .def("submit", [](T& s, const py::object& o, const std::string& p) { return s.submit(o, py::none(), p, py::none()); },
             py::arg("options") = py::none(), py::kw_only(), py::arg("priority") = "normal")
        .def("submit_latest",
             [](T& s, const py::object& g, const py::object& o, const std::string& p) {
                 return s.submit(o, py::none(), p, g);
             },
             py::arg("group"), py::arg("options") = py::none(), py::kw_only(), py::arg("priority") = "interactive")
        .def("submit_manifest",
             [](T& s, const py::object& mf, const py::object& o, const std::string& p) {
                 return s.submit(o, mf, p, py::none());
             },
             py::arg("manifest"), py::arg("options") = py::none(), py::kw_only(), py::arg("priority") = "normal")
        .def("submit_manifest_latest",
             [](T& s, const py::object& g, const py::object& mf, const py::object& o, const std::string& p) {
                 return s.submit(o, mf, p, g);
             },
             py::arg("group"), py::arg("manifest"), py::arg("options") = py::none(), py::kw_only(),
             py::arg("priority") = "interactive")
        .def("export_manifest", [](const T& s, const py::object& o) { return s.export_manifest(o); },
             py::arg("options") = py::none())
        .def("required_source_regions",
             [](const T& s, const py::object& mf, const py::object& o) { return s.required_regions(mf, o); },
             py::arg("manifest"), py::arg("options") = py::none())
        .def("histogram_manifest",
             [](const T& s, const py::object& mf, const py::object& o, unsigned bins, double lower, double upper) {
                 return s.histogram(mf, o, bins, lower, upper);
             },
             py::arg("manifest"), py::arg("options") = py::none(), py::kw_only(), py::arg("bins") = 256u,
             py::arg("lower") = 0.0, py::arg("upper") = 1.0)
        .def("analyze_local_manifest",
             [](const T& s, const py::object& mf, const py::object& cb, const py::object& o, const py::object& radius) {
                 return s.analyze_local(mf, cb, o, radius);
             },
             py::arg("manifest"), py::arg("callback"), py::arg("options") = py::none(), py::kw_only(),
             py::arg("radius") = py::none())
        .def("source_info", [](const T& s) { return s.source_info(); })
        .def("cache_stats", [](const T& s) { return s.cache_stats(); })
        .def("clear_cache", [](T& s) { s.clear_cache(); })
        .def("close", [](T& s) {
            py::gil_scoped_release unlocked;
            s.close();
        });
}

}  // namespace

PYBIND11_MODULE(_rawengine_native, m) {
    m.doc() = "Tiled, non-destructive RAW rendering on MI355X (rawengine_native entry points).";
    static py::exception<RenderCancelled> cancelled(m, "RenderCancelled", PyExc_RuntimeError);
    g_cancelled = cancelled.ptr();
    m.def("icc_available", [] { return false; },
          "Return whether an ICC colour-management backend is compiled into this extension.");
    m.def("create_icc_profile", [](const py::object&, const py::kwargs&) -> py::object {
        raise_py(PyExc_RuntimeError, "ICC profiles require a build with a colour-management engine");
    });
    m.def("render", &render, py::arg("bayer"), py::arg("width"), py::arg("height"), py::arg("options") = py::none(),
          "render(bayer, width, height, options=None) -> (width, height, float32_rgb_bytes)");
    m.def("render_dlpack", &render_dlpack, py::arg("bayer"), py::arg("width"), py::arg("height"),
          py::arg("options") = py::none(),
          "render_dlpack(...) -> (width, height, DLPack capsule of device float32 (height, width, 3))");
    m.def("render_raster", &render_raster, py::arg("rgb"), py::arg("width"), py::arg("height"),
          py::arg("options") = py::none(),
          "render_raster(rgb, width, height, options) -> (width, height, float32_rgb_bytes)");

    py::class_<RenderJob>(m, "RenderJob")
        .def("output_kind", &RenderJob::output_kind)
        .def("result", &RenderJob::result, py::arg("timeout") = py::none())
        .def("done", &RenderJob::done)
        .def("cancel", &RenderJob::cancel)
        .def("progress", &RenderJob::progress);

    bind_session<RawSession>(py::class_<RawSession>(m, "RawSession")
                                 .def(py::init(&raw_session), py::arg("bayer"), py::arg("width"), py::arg("height"),
                                      py::arg("metadata") = py::none(), py::kw_only(),
                                      py::arg("cache_bytes") = size_t(64) << 20, py::arg("workers") = size_t(1),
                                      py::arg("max_pending") = size_t(8), py::arg("demosaic") = py::none(),
                                      py::arg("output_profile") = py::none()));
    bind_session<RasterSession>(py::class_<RasterSession>(m, "RasterSession")
                                    .def(py::init(&raster_session), py::arg("rgb"), py::arg("width"),
                                         py::arg("height"), py::arg("working_space"),
                                         py::arg("row_stride_pixels") = 0u, py::kw_only(),
                                         py::arg("cache_bytes") = size_t(64) << 20, py::arg("workers") = size_t(1),
                                         py::arg("max_pending") = size_t(8)));
}
