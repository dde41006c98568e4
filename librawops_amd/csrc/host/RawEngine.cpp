// rawengine drop-in layer over the gfx950 C ABI.
//
// Node objects are immutable descriptions with the reference's validation and
// level rules (RawEngine.cpp:223-758, 1215-1475; ToneOps.cpp:17-32, 795-844;
// SpatialOps.cpp:159-317).  Rendering walks the chain from the requested node
// to its source, compiles it into one device plan (demosaic + op program +
// mip reduction, or guided filter + op program) and runs it with the source
// resident in HBM.  A node type this library does not know is rendered with
// its own render_level(); the built-in nodes above it then run on the GPU over
// that uploaded tile.
#include "../../../include/rawengine/RawEngine.hpp"
#include "../../../include/rawengine/SpatialOps.hpp"
#include "../../../include/rawengine/ToneOps.hpp"
#include "../../../include/rawops_hip.h"
#include "internal.hpp"
#include "sha256.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

namespace rawengine {
namespace detail {

struct DeviceSource {
    rawops_source* handle = nullptr;
    ~DeviceSource() { rawops_source_destroy(handle); }
};

[[noreturn]] void raise(rawops_status s) {
    const std::string m = rawops_last_error();
    switch (s) {
    case RAWOPS_INVALID_ARGUMENT: throw std::invalid_argument(m);
    case RAWOPS_OUT_OF_RANGE: throw std::out_of_range(m);
    case RAWOPS_DOMAIN_ERROR: throw std::domain_error(m);
    case RAWOPS_LENGTH_ERROR: throw std::length_error(m);
    case RAWOPS_OVERFLOW_ERROR: throw std::overflow_error(m);
    case RAWOPS_CANCELLED: throw RenderCancelled();
    case RAWOPS_LOGIC_ERROR: throw std::logic_error(m);
    default: throw std::runtime_error(m);
    }
}

void hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
// Streams outlive the threads that used them: a finished thread's stream is
// handed to the next thread that asks, so the process creates at most as many
// streams as it ever ran rendering threads at once.
std::mutex stream_pool_mutex;
std::vector<hipStream_t> stream_pool;

struct ThreadStream {
    hipStream_t s = nullptr;
    ~ThreadStream() {
        if (!s) return;
        std::lock_guard lock(stream_pool_mutex);
        stream_pool.push_back(s);
    }
};
thread_local ThreadStream this_thread_stream;
}  // namespace

// Two pinned staging buffers per thread (allocated on first use): chunk i's
// DMA runs while chunk i - 1 is copied on the host into dst.
void copy_to_host(void* dst_v, const void* src, size_t bytes, hipStream_t stream) {
    char* dst = static_cast<char*>(dst_v);
    constexpr size_t kChunk = size_t(16) << 20;
    struct Staging {
        void* pin[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        ~Staging() {
            for (int k = 0; k < 2; ++k) {
                if (pin[k]) (void)hipHostFree(pin[k]);
                if (done[k]) (void)hipEventDestroy(done[k]);
            }
        }
    };
    thread_local Staging st;
    if (!st.pin[0]) {
        for (int k = 0; k < 2; ++k) {
            hip(hipHostMalloc(&st.pin[k], kChunk, hipHostMallocDefault), "hipHostMalloc(staging)");
            hip(hipEventCreateWithFlags(&st.done[k], hipEventDisableTiming), "hipEventCreate");
        }
    }
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    auto len = [&](size_t i) { return std::min(kChunk, bytes - i * kChunk); };
    for (size_t i = 0; i <= chunks; ++i) {
        if (i < chunks) {
            hip(hipMemcpyAsync(st.pin[i & 1], static_cast<const char*>(src) + i * kChunk, len(i),
                               hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(staging)");
            hip(hipEventRecord(st.done[i & 1], stream), "hipEventRecord");
        }
        if (i > 0) {
            const size_t j = i - 1;
            hip(hipEventSynchronize(st.done[j & 1]), "hipEventSynchronize");
            std::memcpy(dst + j * kChunk, st.pin[j & 1], len(j));
        }
    }
}

hipStream_t thread_stream() {
    if (!this_thread_stream.s) {
        {
            std::lock_guard lock(stream_pool_mutex);
            if (!stream_pool.empty()) {
                this_thread_stream.s = stream_pool.back();
                stream_pool.pop_back();
            }
        }
        if (!this_thread_stream.s) hip(hipStreamCreateWithFlags(&this_thread_stream.s, hipStreamNonBlocking),
                                       "hipStreamCreateWithFlags");
    }
    return this_thread_stream.s;
}

DeviceBuffer::DeviceBuffer(size_t bytes, hipStream_t s) : stream(s), pooled(true) {
    // Keep released blocks in the pool instead of trimming it at every
    // synchronisation (once per device).
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
    if (bytes) hip(hipMallocAsync(&p, bytes, s), "hipMallocAsync");
}

DeviceTile::DeviceTile(std::uint32_t w, std::uint32_t h, ImageDescriptor d) : width(w), height(h), descriptor(d) {
    if (floats()) hip(hipMalloc(&ptr, floats() * sizeof(float)), "hipMalloc(tile)");
}

DeviceTile::~DeviceTile() {
    if (ptr) (void)hipFree(ptr);
}

struct Plan {
    rawops_plan* handle = nullptr;
    ~Plan() { rawops_plan_destroy(handle); }
};

struct Op {
    rawops_op_desc desc{};
    std::vector<double> xs, ys;  // curve knots backing desc.knots_x/y
};

// One node of a compiled chain and, for graph nodes, its tile-cache identity.
struct Link {
    const Node* node = nullptr;  // the node itself (a cached graph node is unwrapped)
    TileCache* cache = nullptr;
    Signature signature{};
    bool has_signature = false;
};

// One device launch of a chain: segment 0 reads the source (demosaic + op
// program, or raster reduction + op program, or a foreign node's uploaded
// tile); every further segment is a guided filter + op program over the
// previous segment's device output.
struct Segment {
    const WorkingYGuidedFilterNode* guided = nullptr;
    std::vector<Op> native, reduced;
    bool anchor = false;
    std::unique_ptr<Plan> plan;
    Link last;                 // node whose output this launch produces
    bool cache_boundary = false;
};

struct Chain {
    enum Kind { Raw, Raster, Foreign } kind = Raw;
    const RawImage* raw = nullptr;
    const RasterImage* raster = nullptr;
    const Node* foreign = nullptr;
    uint32_t demosaic = RAWOPS_DEMOSAIC_BILINEAR;
    bool fuse_source = false;  // segment 0 is the bare source: segment 1 filters it directly
    std::vector<Segment> segs;
};

rawops_rect c(Rect r) { return {r.x, r.y, r.width, r.height}; }

Op matrix_op(const std::array<float, 9>& m) {
    Op o;
    o.desc.kind = RAWOPS_OP_MATRIX;
    for (int i = 0; i < 9; ++i) o.desc.v[i] = m[i];
    return o;
}

// Op contributed by one built-in point node; false if `n` is not one.
bool point_ops(const Node* n, std::vector<Op>& out) {
    if (auto* wb = dynamic_cast<const WhiteBalanceNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_GAINS;
        for (int i = 0; i < 3; ++i) o.desc.v[i] = wb->gains()[i];
        out.push_back(o);
    } else if (auto* ex = dynamic_cast<const ExposureNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_SCALE;
        o.desc.v[0] = ex->multiplier();
        out.push_back(o);
    } else if (auto* cw = dynamic_cast<const CameraToWorkingNode*>(n)) {
        out.push_back(matrix_op(cw->matrix()));
    } else if (auto* ws = dynamic_cast<const WorkingSpaceConvertNode*>(n)) {
        if (!ws->identity()) out.push_back(matrix_op(ws->matrix()));
    } else if (auto* ts = dynamic_cast<const WorkingToSrgbNode*>(n)) {
        out.push_back(matrix_op(ts->matrix()));
    } else if (auto* tc = dynamic_cast<const ToneCurveNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_TONE;
        o.desc.v[0] = tc->shoulder();
        o.desc.v[1] = tc->inverse_gamma();
        out.push_back(o);
    } else if (dynamic_cast<const OutputClipNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_CLIP;
        out.push_back(o);
    } else if (dynamic_cast<const SrgbEncodeNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_SRGB_ENCODE;
        out.push_back(o);
    } else if (auto* cv = dynamic_cast<const ExtendedCurvesNode*>(n)) {
        Op o;
        o.desc.kind = RAWOPS_OP_CURVES;
        const auto& s = cv->settings();
        o.desc.knot_offsets[0] = 0;
        for (int k = 0; k < 4; ++k) {
            const auto& knots = k == 0 ? s.master.knots : s.channels[k - 1].knots;
            for (const auto& p : knots) { o.xs.push_back(p.x); o.ys.push_back(p.y); }
            o.desc.knot_offsets[k + 1] = uint32_t(o.xs.size());
        }
        o.desc.cubic = s.interpolation == CurveInterpolation::ShapePreservingCubic ? 1u : 0u;
        out.push_back(std::move(o));
    } else {
        return false;
    }
    return true;
}

std::unique_ptr<Plan> make_plan(uint32_t demosaic, bool anchor, std::vector<Op>& native, std::vector<Op>& reduced) {
    std::vector<rawops_op_desc> a, b;
    for (auto& o : native) {
        if (o.desc.kind == RAWOPS_OP_CURVES) { o.desc.knots_x = o.xs.data(); o.desc.knots_y = o.ys.data(); }
        a.push_back(o.desc);
    }
    for (auto& o : reduced) {
        if (o.desc.kind == RAWOPS_OP_CURVES) { o.desc.knots_x = o.xs.data(); o.desc.knots_y = o.ys.data(); }
        b.push_back(o.desc);
    }
    rawops_plan_desc d{};
    d.demosaic = demosaic;
    d.mip_anchor = anchor ? 1u : 0u;
    d.n_native = uint32_t(a.size());
    d.n_reduced = uint32_t(b.size());
    d.native_ops = a.data();
    d.reduced_ops = b.data();
    auto p = std::make_unique<Plan>();
    check(rawops_plan_create(&d, &p->handle));
    return p;
}

Link link_of(const Node* n) {
    Link l;
    if (auto* cn = dynamic_cast<const CachedNode*>(n)) {
        l.node = &cn->inner();
        l.cache = cn->cache();
        l.signature = cn->signature();
        l.has_signature = true;
    } else {
        l.node = n;
    }
    return l;
}

bool is_source(const Node* n) {
    return dynamic_cast<const RawUnpackNode*>(n) || dynamic_cast<const RasterSourceNode*>(n);
}

std::unique_ptr<Chain> compile(const Node& output) {
    std::vector<Link> links;
    for (const Node* n = &output; n; n = n->input_node()) links.push_back(link_of(n));
    std::reverse(links.begin(), links.end());
    auto ch = std::make_unique<Chain>();
    size_t first = 1;
    if (auto* ru = dynamic_cast<const RawUnpackNode*>(links[0].node)) {
        ch->kind = Chain::Raw;
        ch->raw = &ru->image();
        ch->demosaic = ru->raw_demosaic_identity()->algorithm == "rawengine.menon_base" ? RAWOPS_DEMOSAIC_MENON
                                                                                         : RAWOPS_DEMOSAIC_BILINEAR;
    } else if (auto* rs = dynamic_cast<const RasterSourceNode*>(links[0].node)) {
        ch->kind = Chain::Raster;
        ch->raster = &rs->image();
    } else {
        // Render the last foreign node with its own code; run the rest on the GPU.
        size_t last = 0;
        for (size_t i = 0; i < links.size(); ++i) {
            std::vector<Op> probe;
            const bool known = point_ops(links[i].node, probe) || dynamic_cast<const DeviceNode*>(links[i].node);
            if (!known) last = i;
        }
        ch->kind = Chain::Foreign;
        ch->foreign = links[last].node;
        first = last + 1;
    }
    ch->segs.emplace_back();
    ch->segs.back().last = links[first - 1];
    for (size_t i = first; i < links.size(); ++i) {
        const Link& l = links[i];
        if (auto* gf = dynamic_cast<const WorkingYGuidedFilterNode*>(l.node)) {
            ch->segs.emplace_back();
            ch->segs.back().guided = gf;
            ch->segs.back().last = l;
            continue;
        }
        Segment& s = ch->segs.back();
        auto& list = ch->kind == Chain::Raw && ch->segs.size() == 1 && !s.anchor ? s.native : s.reduced;
        if (!point_ops(l.node, list)) throw std::invalid_argument("node type cannot be compiled into a device plan");
        if (auto* cw = dynamic_cast<const CameraToWorkingNode*>(l.node))
            if (ch->kind == Chain::Raw && ch->segs.size() == 1 && cw->native_input_bounds().width) s.anchor = true;
        s.last = l;
    }
    // A bare raster (or foreign tile) under a guided filter is filtered where
    // it lies: no copy launch.
    const Segment& s0 = ch->segs[0];
    ch->fuse_source = ch->kind != Chain::Raw && ch->segs.size() > 1 && s0.native.empty() && s0.reduced.empty() &&
                      !(ch->kind == Chain::Foreign && s0.last.has_signature && s0.last.cache);
    for (size_t k = 0; k < ch->segs.size(); ++k) {
        Segment& s = ch->segs[k];
        // Source tiles are resident already: a cache entry would only copy them.
        s.cache_boundary = s.last.has_signature && s.last.cache && !is_source(s.last.node) &&
                           !(k == 0 && ch->fuse_source);
        s.plan = make_plan(ch->demosaic, s.anchor, s.native, s.reduced);
    }
    return ch;
}

Rect request_bounds(Rect source, RenderLevel level) {
    rawops_rect out{};
    check(rawops_request_bounds(c(source), level.mip, static_cast<uint32_t>(level.quality), &out));
    return {out.x, out.y, out.width, out.height};
}

WorkingSpace space_of(const ImageDescriptor& d) {
    return d.primaries == ColorPrimaries::Rec2020 ? WorkingSpace::LinearRec2020D65 : WorkingSpace::LinearProPhotoD50;
}

// A device buffer (previous launch's output) viewed as a raster source.
struct WrappedRaster {
    rawops_source* handle = nullptr;
    WrappedRaster(const float* p, uint32_t w, uint32_t h, WorkingSpace ws) {
        check(rawops_source_wrap_raster_device(w, h, w, uint32_t(ws), p, &handle));
    }
    ~WrappedRaster() { rawops_source_destroy(handle); }
};

// Runs the chain over `viewport` (level coordinates) into dst (row stride =
// viewport width).  Segment outputs that are cache boundaries are looked up
// top-down; rendering resumes above the highest hit.  New entries are stored
// only after the launches have finished without an output-boundary error.
void run(const Chain& ch, Rect viewport, RenderLevel level, float* dst, uint32_t* flags, hipStream_t stream) {
    const size_t S = ch.segs.size();
    std::vector<Rect> region(S);
    region[S - 1] = viewport;
    for (size_t k = S - 1; k >= 1; --k) {
        const auto* gf = ch.segs[k].guided;
        region[k - 1] = gf->input_region_level(region[k], request_bounds(gf->native_bounds(), level), level);
    }
    std::vector<std::shared_ptr<const DeviceTile>> out(S);
    std::vector<uint64_t> generation(S, 0);
    size_t start = 0;
    for (size_t k = S; k-- > 0;) {
        const Segment& s = ch.segs[k];
        if (!s.cache_boundary) continue;
        out[k] = CacheAccess::lookup(*s.last.cache, s.last.signature, region[k], level, &generation[k]);
        if (out[k]) { start = k + 1; break; }
    }
    const size_t row_bytes = size_t(viewport.width) * 3 * sizeof(float);
    if (start == S) {
        hip(hipMemcpyAsync(dst, out[S - 1]->ptr, row_bytes * viewport.height, hipMemcpyDeviceToDevice, stream),
            "hipMemcpyAsync(cached tile)");
        return;
    }
    const uint32_t quality = static_cast<uint32_t>(level.quality);
    std::unique_ptr<DeviceSource> foreign_upload;
    Rect foreign_region{};
    for (size_t k = start; k < S; ++k) {
        const Segment& s = ch.segs[k];
        if (k == 0 && ch.fuse_source) continue;
        float* target = dst;
        std::shared_ptr<DeviceTile> fresh;
        if (k + 1 < S) {
            fresh = std::make_shared<DeviceTile>(region[k].width, region[k].height, s.last.node->output_descriptor());
            target = static_cast<float*>(fresh->ptr);
            out[k] = fresh;
        }
        const uint32_t ostride = region[k].width;
        if (k == 0 && ch.kind == Chain::Raw) {
            check(rawops_render(s.plan->handle, ch.raw->device().handle, c(region[0]), level.mip, quality, target,
                                nullptr, ostride, flags, stream));
        } else if (k == 0 && ch.kind == Chain::Raster) {
            check(rawops_guided_filter(ch.raster->device().handle, c(region[0]), level.mip, 0, 1.0, s.plan->handle,
                                       target, ostride, nullptr, 0, flags, stream));
        } else if (k == 0 || (k == 1 && ch.fuse_source && ch.kind == Chain::Foreign)) {
            // Foreign input: its own render at this level, uploaded.
            foreign_region = region[0];
            const Tile t = ch.foreign->render_level(foreign_region, level);
            if (t.bounds.x != foreign_region.x || t.bounds.y != foreign_region.y ||
                t.bounds.width != foreign_region.width || t.bounds.height != foreign_region.height ||
                t.rgb.size() != size_t(foreign_region.width) * foreign_region.height * 3)
                throw std::domain_error("upstream node returned an invalid tile");
            foreign_upload = std::make_unique<DeviceSource>();
            check(rawops_source_create_raster(foreign_region.width, foreign_region.height, 0,
                                              uint32_t(space_of(t.descriptor)), t.rgb.data(), &foreign_upload->handle));
            if (k == 0) {
                check(rawops_guided_filter(foreign_upload->handle, {0, 0, foreign_region.width, foreign_region.height},
                                           0, 0, 1.0, s.plan->handle, target, ostride, nullptr, 0, flags, stream));
            }
        }
        if (k >= 1) {
            const auto* gf = s.guided;
            const Rect in = region[k - 1];
            const rawops_rect v{region[k].x - in.x, region[k].y - in.y, region[k].width, region[k].height};
            if (k == 1 && ch.fuse_source && ch.kind == Chain::Raster) {
                check(rawops_guided_filter(ch.raster->device().handle, c(region[k]), level.mip,
                                           gf->settings().radius, gf->settings().epsilon, s.plan->handle, target,
                                           ostride, nullptr, 0, flags, stream));
            } else if (k == 1 && ch.fuse_source) {
                check(rawops_guided_filter(foreign_upload->handle, v, 0, gf->settings().radius,
                                           gf->settings().epsilon, s.plan->handle, target, ostride, nullptr, 0, flags,
                                           stream));
            } else {
                WrappedRaster src(static_cast<const float*>(out[k - 1]->ptr), in.width, in.height,
                                  space_of(out[k - 1]->descriptor));
                check(rawops_guided_filter(src.handle, v, 0, gf->settings().radius, gf->settings().epsilon,
                                           s.plan->handle, target, ostride, nullptr, 0, flags, stream));
            }
        }
    }
    // Entries enter the caches only once the launches are known to be good.
    bool pending = false;
    for (size_t k = start; k < S; ++k) pending |= ch.segs[k].cache_boundary;
    if (!pending) {
        if (foreign_upload) hip(hipStreamSynchronize(stream), "stream synchronize");
        return;
    }
    if (ch.segs[S - 1].cache_boundary &&
        CacheAccess::admits(*ch.segs[S - 1].last.cache, size_t(viewport.width) * viewport.height * 3)) {
        auto copy = std::make_shared<DeviceTile>(viewport.width, viewport.height,
                                                 ch.segs[S - 1].last.node->output_descriptor());
        hip(hipMemcpyAsync(copy->ptr, dst, row_bytes * viewport.height, hipMemcpyDeviceToDevice, stream),
            "hipMemcpyAsync(cache entry)");
        out[S - 1] = copy;
    }
    hip(hipStreamSynchronize(stream), "stream synchronize");
    if (!flags) return;  // errors of this render are not observable: nothing is cached
    check(rawops_check_flags(flags, stream));
    for (size_t k = start; k < S; ++k) {
        const Segment& s = ch.segs[k];
        if (s.cache_boundary && out[k])
            CacheAccess::store(*s.last.cache, s.last.signature, region[k], level, out[k], generation[k]);
    }
}

void validate_rect(Rect area, Rect r) {
    if (r.x < area.x || r.y < area.y || uint64_t(r.x) + r.width > uint64_t(area.x) + area.width ||
        uint64_t(r.y) + r.height > uint64_t(area.y) + area.height)
        throw std::out_of_range("viewport is outside the active RAW area");
}

const Node* source_of(const Node& n) {
    const Node* s = &n;
    while (s->input_node()) s = s->input_node();
    return s;
}

bool compilable(const Node& n) { return dynamic_cast<const DeviceNode*>(&n) != nullptr; }

bool single_launch_raw_chain(const Node& output) {
    for (const Node* n = &output; n; n = n->input_node()) {
        if (auto* cn = dynamic_cast<const CachedNode*>(n)) {
            if (cn->cache()) return false;
            n = &cn->inner();
        }
        std::vector<Op> probe;
        if (dynamic_cast<const RawUnpackNode*>(n)) return true;
        if (!point_ops(n, probe)) return false;
    }
    return false;
}

bool batch_plan(const Node& output, BatchPlan* out) {
    if (!compilable(output)) return false;
    std::shared_ptr<Chain> ch;
    try {
        ch = compile(output);
    } catch (const std::exception&) {
        return false;
    }
    if (ch->kind != Chain::Raw || ch->segs.size() != 1 || ch->segs[0].cache_boundary) return false;
    out->plan = ch->segs[0].plan->handle;
    out->source = ch->raw->device().handle;
    out->keep = std::move(ch);
    return true;
}

void render_to_device(const Node& output, Rect source_bounds, const RenderRequest& request, float* dst,
                      hipStream_t stream) {
    validate_rect(request_bounds(source_bounds, request.level), request.viewport);
    if (!output.supports_level(request.level)) throw std::invalid_argument("node does not support this render level");
    if (!request.viewport.width || !request.viewport.height) return;
    const auto chain = compile(output);
    DeviceBuffer flags(sizeof(uint32_t), stream);
    hip(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), stream), "hipMemsetAsync(flags)");
    run(*chain, request.viewport, request.level, dst, static_cast<uint32_t*>(flags.p), stream);
    check(rawops_check_flags(static_cast<uint32_t*>(flags.p), stream));
}

}  // namespace detail

using detail::check;

// ------------------------------------------------------------------ images --

RawImage::RawImage(RawMetadata metadata, std::vector<std::uint16_t> samples)
    : metadata_(std::move(metadata)),
      bayer_(std::make_shared<const std::vector<std::uint16_t>>(std::move(samples))) {
    auto& m = metadata_;
    if (!m.width || !m.height || static_cast<int>(m.pattern) < 0 || static_cast<int>(m.pattern) > 3 ||
        m.cfa_phase_x > 1 || m.cfa_phase_y > 1)
        throw std::invalid_argument("invalid RAW dimensions, Bayer pattern, or CFA phase");
    if (!m.row_stride_samples) m.row_stride_samples = m.width;
    if (m.row_stride_samples < m.width) throw std::invalid_argument("RAW row stride is smaller than width");
    if (!m.active_area.x && !m.active_area.y && !m.active_area.width && !m.active_area.height)
        m.active_area = {0, 0, m.width, m.height};
    const auto a = m.active_area;
    if (!a.width || !a.height || uint64_t(a.x) + a.width > m.width || uint64_t(a.y) + a.height > m.height)
        throw std::invalid_argument("RAW active area is outside the sensor");
    for (int s = 0; s < 4; ++s)
        if (m.white_levels[s] <= m.black_levels[s])
            throw std::invalid_argument("RAW white level must exceed black level at every site");
    if (bayer_->size() != size_t(m.row_stride_samples) * m.height)
        throw std::invalid_argument("Bayer sample count does not match stride and height");
}

static RawMetadata uniform(std::uint32_t w, std::uint32_t h, BayerPattern p, std::uint16_t b, std::uint16_t wl) {
    RawMetadata m;
    m.width = w;
    m.height = h;
    m.pattern = p;
    m.black_levels.fill(b);
    m.white_levels.fill(wl);
    return m;
}

RawImage::RawImage(std::uint32_t width, std::uint32_t height, std::vector<std::uint16_t> samples,
                   BayerPattern pattern, std::uint16_t black, std::uint16_t white)
    : RawImage(uniform(width, height, pattern, black, white), std::move(samples)) {}

const detail::DeviceSource& RawImage::device() const {
    std::call_once(*upload_once_, [&] {
        auto src = std::make_shared<detail::DeviceSource>();
        rawops_raw_metadata m{};
        m.width = metadata_.width; m.height = metadata_.height; m.row_stride_samples = metadata_.row_stride_samples;
        m.pattern = uint32_t(metadata_.pattern);
        m.cfa_phase_x = metadata_.cfa_phase_x; m.cfa_phase_y = metadata_.cfa_phase_y;
        m.active_area = detail::c(metadata_.active_area);
        for (int i = 0; i < 4; ++i) { m.black_levels[i] = metadata_.black_levels[i]; m.white_levels[i] = metadata_.white_levels[i]; }
        check(rawops_source_create_raw(&m, bayer_->data(), bayer_->size(), &src->handle));
        *device_ = src;
    });
    return **device_;
}

RasterImage::RasterImage(RasterMetadata metadata, std::vector<float> pixels)
    : metadata_(metadata), pixels_(std::make_shared<const std::vector<float>>(std::move(pixels))) {
    if (!metadata_.width || !metadata_.height ||
        (metadata_.working_space != WorkingSpace::LinearProPhotoD50 &&
         metadata_.working_space != WorkingSpace::LinearRec2020D65))
        throw std::invalid_argument("invalid scene-linear raster metadata");
    if (!metadata_.row_stride_pixels) metadata_.row_stride_pixels = metadata_.width;
    if (metadata_.row_stride_pixels < metadata_.width ||
        pixels_->size() != size_t(metadata_.row_stride_pixels) * metadata_.height * 3)
        throw std::invalid_argument("raster float count does not match stride and height");
    if (!std::all_of(pixels_->begin(), pixels_->end(), [](float v) { return std::isfinite(v); }))
        throw std::invalid_argument("scene-linear raster pixels must be finite");
}

const detail::DeviceSource& RasterImage::device() const {
    std::call_once(*upload_once_, [&] {
        auto src = std::make_shared<detail::DeviceSource>();
        check(rawops_source_create_raster(metadata_.width, metadata_.height, metadata_.row_stride_pixels,
                                          uint32_t(metadata_.working_space), pixels_->data(), &src->handle));
        *device_ = src;
    });
    return **device_;
}

void validate_raw_demosaic(const RawDemosaicIdentity& identity) {
    if ((identity.algorithm != "rawengine.bilinear" && identity.algorithm != "rawengine.menon_base") ||
        identity.processing_version != 1)
        throw std::invalid_argument("unsupported RAW demosaic algorithm or processing version");
}

// ------------------------------------------------------------------- nodes --

Tile DeviceNode::render(Rect bounds) const { return render_level(bounds, {}); }

Tile DeviceNode::render_level(Rect bounds, RenderLevel level) const {
    if (!supports_level(level)) throw std::invalid_argument("node does not support this render level");
    const Node* src = detail::source_of(*this);
    const auto sb = src->source_bounds();
    Rect source = sb ? *sb : bounds;
    if (!sb && level.mip) throw std::invalid_argument("node has no source bounds for reduced levels");
    const uint32_t tile = std::max<uint32_t>(1, std::max(bounds.width, bounds.height));
    return Renderer{}.render_image(*this, source, RenderRequest{bounds, tile, level});
}

RawUnpackNode::RawUnpackNode(RawImage image) : RawUnpackNode(std::move(image), RawDemosaicIdentity{}) {}

RawUnpackNode::RawUnpackNode(RawImage image, RawDemosaicIdentity demosaic)
    : image_(std::move(image)), demosaic_(std::move(demosaic)) {
    validate_raw_demosaic(demosaic_);
}

Rect RawUnpackNode::input_region(Rect output, Rect source_bounds) const {
    const auto a = image_.metadata().active_area;
    if (source_bounds.x != a.x || source_bounds.y != a.y || source_bounds.width != a.width ||
        source_bounds.height != a.height)
        throw std::invalid_argument("RAW source bounds differ from active area");
    detail::validate_rect(a, output);
    const uint32_t r = demosaic_.algorithm == "rawengine.menon_base" && a.width > 1 && a.height > 1 ? 6u : 1u;
    const auto left = std::max<int64_t>(a.x, int64_t(output.x) - r);
    const auto top = std::max<int64_t>(a.y, int64_t(output.y) - r);
    const auto right = std::min<uint64_t>(uint64_t(a.x) + a.width, uint64_t(output.x) + output.width + r);
    const auto bottom = std::min<uint64_t>(uint64_t(a.y) + a.height, uint64_t(output.y) + output.height + r);
    return {uint32_t(left), uint32_t(top), uint32_t(right - left), uint32_t(bottom - top)};
}

RasterSourceNode::RasterSourceNode(RasterImage image) : image_(std::move(image)) {}

WhiteBalanceNode::WhiteBalanceNode(std::shared_ptr<const Node> input, float red, float green, float blue)
    : input_(std::move(input)), gains_{red, green, blue} {
    if (!input_) throw std::invalid_argument("white balance input is null");
    if (input_->output_descriptor() != ImageDescriptor::camera_linear())
        throw std::invalid_argument("white balance requires camera-linear RGB input");
    for (float g : gains_)
        if (!std::isfinite(g) || g <= 0.0f || g > 65536.0f)
            throw std::invalid_argument("white balance gains must be finite and in (0, 65536]");
}

ExposureNode::ExposureNode(std::shared_ptr<const Node> input, float stops) : input_(std::move(input)) {
    if (!input_ || !std::isfinite(stops) || stops < -32.0f || stops > 32.0f)
        throw std::invalid_argument("exposure stops must be finite and in [-32, 32]");
    check(rawops_exposure_multiplier(stops, &multiplier_));
    descriptor_ = input_->output_descriptor();
    if (descriptor_ != ImageDescriptor::camera_linear() &&
        descriptor_ != ImageDescriptor::scene_linear(WorkingSpace::LinearProPhotoD50) &&
        descriptor_ != ImageDescriptor::scene_linear(WorkingSpace::LinearRec2020D65))
        throw std::invalid_argument("exposure requires linear RGB input");
}

bool ExposureNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return true;
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview &&
           descriptor_.domain == PixelDomain::SceneLinearRGB && input_->supports_level(level);
}

CameraToWorkingNode::CameraToWorkingNode(std::shared_ptr<const Node> input, CameraColorTransform transform,
                                         Rect native_input_bounds)
    : input_(std::move(input)), descriptor_(ImageDescriptor::scene_linear(transform.target)),
      native_input_bounds_(native_input_bounds) {
    if (!input_ || input_->output_descriptor() != ImageDescriptor::camera_linear())
        throw std::invalid_argument("color transform requires camera-linear RGB input");
    const auto b = native_input_bounds_;
    if ((b.width == 0) != (b.height == 0) || (!b.width && (b.x || b.y)) ||
        uint64_t(b.x) + b.width > uint64_t(std::numeric_limits<uint32_t>::max()) + 1 ||
        uint64_t(b.y) + b.height > uint64_t(std::numeric_limits<uint32_t>::max()) + 1)
        throw std::invalid_argument("invalid camera preview input bounds");
    check(rawops_camera_to_working_matrix(transform.camera_to_xyz_d50.data(), uint32_t(transform.target),
                                          matrix_.data()));
}

bool CameraToWorkingNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return input_->supports_level(level);
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview &&
           native_input_bounds_.width && native_input_bounds_.height && input_->supports_level({});
}

RenderLevel CameraToWorkingNode::input_level(RenderLevel level) const {
    if (!supports_level(level)) throw std::invalid_argument("camera calibration has no mapping for this level");
    return {};
}

Rect CameraToWorkingNode::input_region_level(Rect output, Rect input_bounds, RenderLevel level) const {
    if (!supports_level(level)) throw std::invalid_argument("camera calibration has no mapping for this level");
    if (!level.mip) return output;
    const auto a = native_input_bounds_;
    if (input_bounds.x != a.x || input_bounds.y != a.y || input_bounds.width != a.width ||
        input_bounds.height != a.height)
        throw std::invalid_argument("camera preview input bounds differ from construction");
    detail::validate_rect(detail::request_bounds(a, level), output);
    const auto s = 1u << level.mip;
    const auto left = std::min<uint64_t>(uint64_t(output.x) * s, a.width);
    const auto top = std::min<uint64_t>(uint64_t(output.y) * s, a.height);
    const auto right = std::min<uint64_t>((uint64_t(output.x) + output.width) * s, a.width);
    const auto bottom = std::min<uint64_t>((uint64_t(output.y) + output.height) * s, a.height);
    return {uint32_t(a.x + left), uint32_t(a.y + top), uint32_t(right - left), uint32_t(bottom - top)};
}

static bool scene_linear(ImageDescriptor d, WorkingSpace* space) {
    if (d == ImageDescriptor::scene_linear(WorkingSpace::LinearProPhotoD50)) { *space = WorkingSpace::LinearProPhotoD50; return true; }
    if (d == ImageDescriptor::scene_linear(WorkingSpace::LinearRec2020D65)) { *space = WorkingSpace::LinearRec2020D65; return true; }
    return false;
}

WorkingSpaceConvertNode::WorkingSpaceConvertNode(std::shared_ptr<const Node> input, WorkingSpace target)
    : input_(std::move(input)), descriptor_(ImageDescriptor::scene_linear(target)) {
    if (!input_) throw std::invalid_argument("working-space conversion input is null");
    WorkingSpace source;
    if (!scene_linear(input_->output_descriptor(), &source))
        throw std::invalid_argument("working-space conversion requires declared scene-linear RGB");
    identity_ = source == target;
    check(rawops_working_to_working_matrix(uint32_t(source), uint32_t(target), matrix_.data()));
}

bool WorkingSpaceConvertNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return true;
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview && input_->supports_level(level);
}

WorkingToSrgbNode::WorkingToSrgbNode(std::shared_ptr<const Node> input) : input_(std::move(input)) {
    if (!input_) throw std::invalid_argument("working-to-sRGB input is null");
    WorkingSpace source;
    if (!scene_linear(input_->output_descriptor(), &source))
        throw std::invalid_argument("sRGB conversion requires declared scene-linear working RGB");
    check(rawops_working_to_srgb_matrix(uint32_t(source), matrix_.data()));
}

bool WorkingToSrgbNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return true;
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview && input_->supports_level(level);
}

ToneCurveNode::ToneCurveNode(std::shared_ptr<const Node> input, float shoulder, float gamma)
    : input_(std::move(input)), shoulder_(shoulder), inverse_gamma_(1.0f / gamma) {
    if (!input_ || !std::isfinite(shoulder) || shoulder <= 0.0f || !std::isfinite(gamma) || gamma <= 0.0f ||
        !std::isfinite(inverse_gamma_))
        throw std::invalid_argument("tone shoulder and gamma must be finite and positive");
    const auto source = input_->output_descriptor();
    WorkingSpace ws;
    if (source != ImageDescriptor::camera_linear() && !scene_linear(source, &ws) &&
        source != ImageDescriptor::linear_srgb())
        throw std::invalid_argument("tone curve requires linear RGB input");
    if (source == ImageDescriptor::linear_srgb()) {
        descriptor_ = ImageDescriptor::display_linear_srgb();
    } else {
        descriptor_ = ImageDescriptor::tone_mapped();
        descriptor_.primaries = source.primaries;
        descriptor_.white_point = source.white_point;
    }
}

bool ToneCurveNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return true;
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview &&
           input_->output_descriptor() == ImageDescriptor::linear_srgb() && input_->supports_level(level);
}

OutputClipNode::OutputClipNode(std::shared_ptr<const Node> input) : input_(std::move(input)) {
    if (!input_ || input_->output_descriptor().domain != PixelDomain::ToneMappedUnmanagedRGB)
        throw std::invalid_argument("output clip requires tone-mapped RGB input");
    descriptor_ = input_->output_descriptor();
    descriptor_.domain = PixelDomain::BoundedUnmanagedRGB;
}

SrgbEncodeNode::SrgbEncodeNode(std::shared_ptr<const Node> input) : input_(std::move(input)) {
    if (!input_) throw std::invalid_argument("sRGB encode input is null");
    if (input_->output_descriptor() != ImageDescriptor::display_linear_srgb())
        throw std::invalid_argument("sRGB encode requires display-linear sRGB input");
}

bool SrgbEncodeNode::supports_level(RenderLevel level) const noexcept {
    if (level.mip == 0 && level.quality == RenderQuality::Final) return true;
    return level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview && input_->supports_level(level);
}

void validate_extended_curves_settings(const ExtendedCurvesSettings& settings) {
    if (settings.interpolation != CurveInterpolation::Linear &&
        settings.interpolation != CurveInterpolation::ShapePreservingCubic)
        throw std::invalid_argument("extended curves interpolation is unsupported");
    auto validate = [](const CurveKnots& c) {
        if (c.knots.size() < 2 || c.knots.size() > 256)
            throw std::invalid_argument("curve knot count exceeds selected operation bounds");
        for (size_t i = 0; i < c.knots.size(); ++i) {
            const auto& p = c.knots[i];
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || std::abs(p.x) > 65536 || std::abs(p.y) > 65536)
                throw std::invalid_argument("curve coordinates must be finite with magnitude <=65536");
            if (i) {
                const auto& q = c.knots[i - 1];
                if (!(p.x > q.x)) throw std::invalid_argument("curve x coordinates must increase strictly");
                const double slope = (p.y - q.y) / (p.x - q.x);
                if (!std::isfinite(slope) || std::abs(slope) > 65536)
                    throw std::invalid_argument("curve slope must be finite with magnitude <=65536");
            }
        }
    };
    validate(settings.master);
    for (const auto& c : settings.channels) validate(c);
}

ExtendedCurvesNode::ExtendedCurvesNode(std::shared_ptr<const Node> input, ExtendedCurvesSettings settings)
    : input_(std::move(input)), settings_(std::move(settings)) {
    validate_extended_curves_settings(settings_);
    WorkingSpace ws;
    if (!input_ || !scene_linear(input_->output_descriptor(), &ws))
        throw std::invalid_argument("extended curves/levels require scene-linear working RGB");
}

bool ExtendedCurvesNode::supports_level(RenderLevel level) const noexcept {
    return ((level.mip == 0 && (level.quality == RenderQuality::Final || level.quality == RenderQuality::Preview)) ||
            (level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview)) &&
           input_->supports_level(level);
}

Rect ExtendedCurvesNode::input_region_level(Rect r, Rect, RenderLevel level) const {
    if (!supports_level(level)) throw std::invalid_argument("extended curves do not support this render level");
    return r;
}

void validate_working_y_guided_filter_settings(const WorkingYGuidedFilterSettings& s) {
    if (s.radius > 8 || !std::isfinite(s.epsilon) || s.epsilon < 0x1p-24 || s.epsilon > 65536.0)
        throw std::invalid_argument("guided filter requires radius 0..8 and finite epsilon [2^-24,65536]");
}

Rect working_y_guided_filter_region(Rect output, Rect image, std::uint32_t radius) {
    if (radius > 8) throw std::invalid_argument("guided filter radius must be 0..8");
    if (output.x < image.x || output.y < image.y || uint64_t(output.x) + output.width > uint64_t(image.x) + image.width ||
        uint64_t(output.y) + output.height > uint64_t(image.y) + image.height)
        throw std::out_of_range("spatial viewport is outside image bounds");
    const uint32_t r = 2 * radius;
    const auto left = std::max<int64_t>(image.x, int64_t(output.x) - r);
    const auto top = std::max<int64_t>(image.y, int64_t(output.y) - r);
    const auto right = std::min<uint64_t>(uint64_t(image.x) + image.width, uint64_t(output.x) + output.width + r);
    const auto bottom = std::min<uint64_t>(uint64_t(image.y) + image.height, uint64_t(output.y) + output.height + r);
    return {uint32_t(left), uint32_t(top), uint32_t(right - left), uint32_t(bottom - top)};
}

WorkingYGuidedFilterNode::WorkingYGuidedFilterNode(std::shared_ptr<const Node> input, Rect native_bounds,
                                                   WorkingYGuidedFilterSettings settings)
    : input_(std::move(input)), native_bounds_(native_bounds), settings_(settings) {
    if (!native_bounds_.width || !native_bounds_.height ||
        uint64_t(native_bounds_.x) + native_bounds_.width > (uint64_t{1} << 32) ||
        uint64_t(native_bounds_.y) + native_bounds_.height > (uint64_t{1} << 32))
        throw std::invalid_argument("spatial operation requires nonempty addressable bounds");
    validate_working_y_guided_filter_settings(settings_);
    if (!input_) throw std::invalid_argument("guided filter requires input");
    WorkingSpace ws;
    descriptor_ = input_->output_descriptor();
    if (!scene_linear(descriptor_, &ws)) throw std::invalid_argument("guided filter requires scene-linear working RGB");
}

bool WorkingYGuidedFilterNode::supports_level(RenderLevel level) const noexcept {
    return ((level.mip == 0 && (level.quality == RenderQuality::Final || level.quality == RenderQuality::Preview)) ||
            (level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview)) &&
           input_->supports_level(level);
}

Rect WorkingYGuidedFilterNode::input_region(Rect output, Rect source_bounds) const {
    return input_region_level(output, source_bounds, {});
}

Rect WorkingYGuidedFilterNode::input_region_level(Rect output, Rect source_bounds, RenderLevel level) const {
    const auto image = detail::request_bounds(native_bounds_, level);
    if (source_bounds.x != image.x || source_bounds.y != image.y || source_bounds.width != image.width ||
        source_bounds.height != image.height || !supports_level(level))
        throw std::invalid_argument("guided filter source bounds or level differ");
    return working_y_guided_filter_region(output, image, settings_.radius);
}

// ------------------------------------------------------------------- graph --

ImageGraph::ImageGraph(RawImage image, GraphRecipe recipe, std::shared_ptr<const IccDisplayTransform> transform)
    : ImageGraph(std::move(image), recipe, RawDemosaicIdentity{}, false, std::move(transform)) {}

ImageGraph::ImageGraph(RawImage image, GraphRecipe recipe, RawDemosaicIdentity demosaic, bool linear_output,
                       std::shared_ptr<const IccDisplayTransform> transform)
    : raw_image_(std::move(image)), source_bounds_(raw_image_->metadata().active_area), recipe_(recipe) {
    std::shared_ptr<const Node> n = std::make_shared<RawUnpackNode>(*raw_image_, std::move(demosaic));
    n = std::make_shared<WhiteBalanceNode>(n, recipe.red_gain, recipe.green_gain, recipe.blue_gain);
    n = std::make_shared<ExposureNode>(n, recipe.exposure_stops);
    if (recipe.camera_color) n = std::make_shared<CameraToWorkingNode>(n, *recipe.camera_color, source_bounds_);
    finish(n, linear_output, std::move(transform));
}

ImageGraph::ImageGraph(RasterImage image, GraphRecipe recipe, std::shared_ptr<const IccDisplayTransform> transform)
    : ImageGraph(std::make_shared<RasterSourceNode>(image), Rect{0, 0, image.width(), image.height()}, recipe,
                 std::move(transform)) {}

ImageGraph::ImageGraph(std::shared_ptr<const Node> source, Rect bounds, GraphRecipe recipe,
                       std::shared_ptr<const IccDisplayTransform> transform)
    : source_bounds_(bounds), recipe_(recipe) {
    if (!source || !bounds.width || !bounds.height)
        throw std::invalid_argument("scene-linear source and valid bounds are required");
    WorkingSpace ws;
    if (!scene_linear(source->output_descriptor(), &ws))
        throw std::invalid_argument("graph source must output a supported scene-linear space");
    if (recipe.red_gain != 1.0f || recipe.green_gain != 1.0f || recipe.blue_gain != 1.0f || recipe.camera_color)
        throw std::invalid_argument("raster input does not accept RAW calibration controls");
    finish(std::make_shared<ExposureNode>(source, recipe.exposure_stops), false, std::move(transform));
}

void ImageGraph::finish(std::shared_ptr<const Node> linear, bool linear_output,
                        std::shared_ptr<const IccDisplayTransform> transform) {
    if (recipe_.output_mode != OutputMode::LegacyBounded && recipe_.output_mode != OutputMode::SrgbPreview &&
        recipe_.output_mode != OutputMode::IccDisplay)
        throw std::invalid_argument("unknown output mode");
    if (linear_output) {
        output_ = linear;
        return;
    }
    if (recipe_.output_mode == OutputMode::SrgbPreview || recipe_.output_mode == OutputMode::IccDisplay)
        linear = std::make_shared<WorkingToSrgbNode>(linear);
    auto tone = std::make_shared<ToneCurveNode>(linear, recipe_.tone_shoulder, recipe_.tone_gamma);
    if (recipe_.output_mode == OutputMode::IccDisplay)
        output_ = std::make_shared<IccDisplayNode>(tone, std::move(transform));
    else
        output_ = recipe_.output_mode == OutputMode::SrgbPreview
                      ? std::static_pointer_cast<const Node>(std::make_shared<SrgbEncodeNode>(tone))
                      : std::static_pointer_cast<const Node>(std::make_shared<OutputClipNode>(tone));
}

// ------------------------------------------------------------- ICC display --

IccDisplayNode::IccDisplayNode(std::shared_ptr<const Node> input, std::shared_ptr<const IccDisplayTransform> transform)
    : input_(std::move(input)), transform_(std::move(transform)) {
    if (!input_ || !transform_ || input_->output_descriptor() != ImageDescriptor::display_linear_srgb())
        throw std::invalid_argument("ICC display requires display-linear sRGB and a transform");
    const auto digest = transform_->profile_sha256();
    if (std::all_of(digest.begin(), digest.end(), [](std::uint8_t v) { return v == 0; }))
        throw std::invalid_argument("ICC output profile digest must be nonzero");
    descriptor_ = ImageDescriptor::icc_display(digest);
}

Tile IccDisplayNode::render(Rect r) const {
    // The display-linear input renders on the device (one launch for the chain).
    Tile tile = input_->render(r);
    if (tile.bounds.x != r.x || tile.bounds.y != r.y || tile.bounds.width != r.width ||
        tile.bounds.height != r.height || tile.rgb.size() != size_t(r.width) * r.height * 3 ||
        tile.descriptor != input_->output_descriptor())
        throw std::domain_error("upstream node returned an invalid tile");
    if (!std::all_of(tile.rgb.begin(), tile.rgb.end(), [](float v) { return std::isfinite(v); }))
        throw std::domain_error("non-finite value at ICC input boundary");
    if (!tile.rgb.empty()) transform_->apply(tile.rgb.data(), tile.rgb.size() / 3);
    if (!std::all_of(tile.rgb.begin(), tile.rgb.end(),
                     [](float v) { return std::isfinite(v) && v >= 0.0f && v <= 1.0f; }))
        throw std::domain_error("ICC display transform returned unbounded or non-finite RGB");
    tile.descriptor = descriptor_;
    return tile;
}

// ------------------------------------------------------------ fingerprints --

namespace {

void put_u8(detail::Sha256& h, uint8_t v) { h.update(&v, 1); }
void put_u16(detail::Sha256& h, uint16_t v) {
    const uint8_t b[2] = {uint8_t(v), uint8_t(v >> 8)};
    h.update(b, 2);
}
void put_u32(detail::Sha256& h, uint32_t v) {
    const uint8_t b[4] = {uint8_t(v), uint8_t(v >> 8), uint8_t(v >> 16), uint8_t(v >> 24)};
    h.update(b, 4);
}
void put_tag(detail::Sha256& h, const char* tag) { h.update(tag, std::strlen(tag) + 1); }

// Little-endian bytes of a row of samples, staged through a stack block.
template <class T> void put_row(detail::Sha256& h, const T* p, size_t n) {
    static_assert(sizeof(T) == 2 || sizeof(T) == 4);
    uint8_t block[4096];
    constexpr size_t per = sizeof(block) / sizeof(T);
    while (n) {
        const size_t m = std::min(n, per);
        for (size_t i = 0; i < m; ++i) {
            uint32_t bits;
            if constexpr (sizeof(T) == 2) bits = p[i];
            else bits = std::bit_cast<uint32_t>(p[i]);
            for (size_t b = 0; b < sizeof(T); ++b) block[i * sizeof(T) + b] = uint8_t(bits >> (8 * b));
        }
        h.update(block, m * sizeof(T));
        p += m;
        n -= m;
    }
}

}  // namespace

std::array<std::uint8_t, 32> RawImage::fingerprint() const {
    std::call_once(*fingerprint_once_, [&] {
        const auto& m = metadata_;
        detail::Sha256 h;
        put_tag(h, "librawops.source.raw.u16.v1");
        put_u32(h, m.width);
        put_u32(h, m.height);
        put_u8(h, uint8_t(m.pattern));
        put_u8(h, m.cfa_phase_x);
        put_u8(h, m.cfa_phase_y);
        put_u32(h, m.active_area.x);
        put_u32(h, m.active_area.y);
        put_u32(h, m.active_area.width);
        put_u32(h, m.active_area.height);
        for (auto v : m.black_levels) put_u16(h, v);
        for (auto v : m.white_levels) put_u16(h, v);
        for (uint32_t y = m.active_area.y; y < m.active_area.y + m.active_area.height; ++y)
            put_row(h, bayer_->data() + size_t(y) * m.row_stride_samples + m.active_area.x, m.active_area.width);
        *fingerprint_ = h.finish();
    });
    return *fingerprint_;
}

std::array<std::uint8_t, 32> RasterImage::fingerprint() const {
    std::call_once(*fingerprint_once_, [&] {
        const auto& m = metadata_;
        detail::Sha256 h;
        put_tag(h, "librawops.source.raster.f32.v1");
        put_u32(h, m.width);
        put_u32(h, m.height);
        put_u8(h, uint8_t(m.working_space));
        for (uint32_t y = 0; y < m.height; ++y)
            put_row(h, pixels_->data() + size_t(y) * m.row_stride_pixels * 3, size_t(m.width) * 3);
        *fingerprint_ = h.finish();
    });
    return *fingerprint_;
}

std::array<std::uint8_t, 32> fingerprint_raw_source(const RawImage& image) { return image.fingerprint(); }
std::array<std::uint8_t, 32> fingerprint_raster_source(const RasterImage& image) { return image.fingerprint(); }

const RawImage& ImageGraph::image() const {
    if (!raw_image_) throw std::logic_error("graph has a raster source, not a RAW source");
    return *raw_image_;
}

// ---------------------------------------------------------------- renderer --

void Renderer::render_tiles(const ImageGraph& graph, RenderRequest request, const TileCallback& callback,
                            const CancellationToken* cancellation) const {
    render_tiles(graph.output(), graph.source_bounds(), request, callback, cancellation);
}

void Renderer::render_tiles(const Node& output, Rect source_bounds, RenderRequest request,
                            const TileCallback& callback, const CancellationToken* cancellation) const {
    detail::validate_rect(detail::request_bounds(source_bounds, request.level), request.viewport);
    if (!output.supports_level(request.level)) throw std::invalid_argument("node does not support this render level");
    if (!callback || !request.tile_size) throw std::invalid_argument("callback and tile size are required");
    if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
    const auto v = request.viewport;
    const uint64_t right = uint64_t(v.x) + v.width, bottom = uint64_t(v.y) + v.height;
    if (!detail::compilable(output)) {
        // A foreign output node renders through its own code, tile by tile.
        for (uint64_t y = v.y; y < bottom; y += request.tile_size)
            for (uint64_t x = v.x; x < right; x += request.tile_size) {
                if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
                const Rect r{uint32_t(x), uint32_t(y), uint32_t(std::min<uint64_t>(request.tile_size, right - x)),
                             uint32_t(std::min<uint64_t>(request.tile_size, bottom - y))};
                callback(output.render_level(r, request.level));
            }
        return;
    }
    if (!v.width || !v.height) return;
    // One device render per row of tiles: the band is produced by one launch,
    // copied to pinned host memory, then handed out tile by tile.
    const auto chain = detail::compile(output);
    const hipStream_t stream = detail::thread_stream();
    const uint32_t band_h = std::min<uint32_t>(request.tile_size, v.height);
    detail::DeviceBuffer dev(size_t(v.width) * band_h * 3 * sizeof(float), stream);
    detail::DeviceBuffer flags(sizeof(uint32_t), stream);
    float* host = nullptr;
    detail::hip(hipHostMalloc(reinterpret_cast<void**>(&host), size_t(v.width) * band_h * 3 * sizeof(float),
                              hipHostMallocDefault), "hipHostMalloc");
    struct Free { float* p; ~Free() { (void)hipHostFree(p); } } free_host{host};
    detail::hip(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), stream), "hipMemsetAsync");
    const auto desc = output.output_descriptor();
    for (uint64_t y = v.y; y < bottom; y += request.tile_size) {
        if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
        const Rect band{v.x, uint32_t(y), v.width, uint32_t(std::min<uint64_t>(request.tile_size, bottom - y))};
        detail::run(*chain, band, request.level, static_cast<float*>(dev.p), static_cast<uint32_t*>(flags.p), stream);
        detail::hip(hipMemcpyAsync(host, dev.p, size_t(band.width) * band.height * 3 * sizeof(float),
                                   hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
        check(rawops_check_flags(static_cast<uint32_t*>(flags.p), stream));
        for (uint64_t x = v.x; x < right; x += request.tile_size) {
            if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
            Tile t{Rect{uint32_t(x), band.y, uint32_t(std::min<uint64_t>(request.tile_size, right - x)), band.height},
                   {}, desc};
            t.rgb.resize(size_t(t.bounds.width) * t.bounds.height * 3);
            for (uint32_t row = 0; row < t.bounds.height; ++row)
                std::memcpy(t.rgb.data() + size_t(row) * t.bounds.width * 3,
                            host + (size_t(row) * band.width + (t.bounds.x - v.x)) * 3,
                            size_t(t.bounds.width) * 3 * sizeof(float));
            callback(t);
        }
    }
}

void Renderer::render_tiles(const ImageGraph& graph, Rect viewport, const TileCallback& callback,
                            std::uint32_t tile_size, const CancellationToken* cancellation) const {
    render_tiles(graph.output(), graph.source_bounds(), RenderRequest{viewport, tile_size, {}}, callback, cancellation);
}

void Renderer::render_tiles(const Node& output, Rect source_bounds, Rect viewport, const TileCallback& callback,
                            std::uint32_t tile_size, const CancellationToken* cancellation) const {
    render_tiles(output, source_bounds, RenderRequest{viewport, tile_size, {}}, callback, cancellation);
}

Tile Renderer::render_image(const ImageGraph& graph, RenderRequest request,
                            const CancellationToken* cancellation) const {
    return render_image(graph.output(), graph.source_bounds(), request, cancellation);
}

Tile Renderer::render_image(const Node& output, Rect source_bounds, RenderRequest request,
                            const CancellationToken* cancellation) const {
    detail::validate_rect(detail::request_bounds(source_bounds, request.level), request.viewport);
    if (!output.supports_level(request.level)) throw std::invalid_argument("node does not support this render level");
    if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
    const auto v = request.viewport;
    Tile image{v, std::vector<float>(size_t(v.width) * v.height * 3), output.output_descriptor()};
    if (!v.width || !v.height) return image;
    if (!detail::compilable(output)) {
        render_tiles(output, source_bounds, request, [&](const Tile& t) {
            for (uint32_t row = 0; row < t.bounds.height; ++row)
                std::memcpy(image.rgb.data() + (size_t(t.bounds.y - v.y + row) * v.width + t.bounds.x - v.x) * 3,
                            t.rgb.data() + size_t(row) * t.bounds.width * 3, size_t(t.bounds.width) * 3 * sizeof(float));
        }, cancellation);
        return image;
    }
    const auto chain = detail::compile(output);
    const hipStream_t stream = detail::thread_stream();
    detail::DeviceBuffer dev(image.rgb.size() * sizeof(float), stream);
    detail::DeviceBuffer flags(sizeof(uint32_t), stream);
    detail::hip(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), stream), "hipMemsetAsync");
    detail::run(*chain, v, request.level, static_cast<float*>(dev.p), static_cast<uint32_t*>(flags.p), stream);
    detail::copy_to_host(image.rgb.data(), dev.p, image.rgb.size() * sizeof(float), stream);
    check(rawops_check_flags(static_cast<uint32_t*>(flags.p), stream));
    return image;
}

Tile Renderer::render_image(const ImageGraph& graph, Rect viewport, std::uint32_t tile_size,
                            const CancellationToken* cancellation) const {
    return render_image(graph.output(), graph.source_bounds(), RenderRequest{viewport, tile_size, {}}, cancellation);
}

Tile Renderer::render_image(const Node& output, Rect source_bounds, Rect viewport, std::uint32_t tile_size,
                            const CancellationToken* cancellation) const {
    return render_image(output, source_bounds, RenderRequest{viewport, tile_size, {}}, cancellation);
}

std::vector<float> Renderer::render_roi(const ImageGraph& graph, Rect viewport, std::uint32_t tile_size,
                                        const CancellationToken* cancellation) const {
    return std::move(render_image(graph, viewport, tile_size, cancellation).rgb);
}

void Renderer::render_device(const Node& output, Rect source_bounds, RenderRequest request, float* device_rgb,
                             void* stream) const {
    detail::validate_rect(detail::request_bounds(source_bounds, request.level), request.viewport);
    if (!output.supports_level(request.level)) throw std::invalid_argument("node does not support this render level");
    const auto chain = detail::compile(output);
    detail::run(*chain, request.viewport, request.level, device_rgb, nullptr, static_cast<hipStream_t>(stream));
}

}  // namespace rawengine

// ---------------------------------------------------------------- analysis --

#include "../../../include/rawengine/ImageAnalysis.hpp"

namespace rawengine {

namespace {

float from_order_key(uint32_t key) {
    const uint32_t bits = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
    float v;
    std::memcpy(&v, &bits, sizeof v);
    return v;
}

}  // namespace

RgbHistogram histogram_rgb(const Node& output, Rect source_bounds, const RenderRequest& request,
                           const HistogramOptions& options, const CancellationToken* cancellation) {
    const double span = options.upper - options.lower;
    if (options.bins == 0 || options.bins > 65536 || !std::isfinite(options.lower) ||
        !std::isfinite(options.upper) || !(span > 0.0) || !std::isfinite(span))
        throw std::invalid_argument("histogram requires 1..65536 bins and a finite positive range");
    auto check_cancelled = [&] {
        if (cancellation && cancellation->is_cancelled()) throw RenderCancelled();
    };
    check_cancelled();
    RgbHistogram result;
    result.request = request;
    result.options = options;
    result.descriptor = output.output_descriptor();
    if (result.descriptor.format != PixelFormat::RGBFloat32)
        throw std::invalid_argument("histogram requires RGBFloat32 output");
    const Rect v = request.viewport;
    const uint64_t right = uint64_t(v.x) + v.width, bottom = uint64_t(v.y) + v.height;
    if (!v.width || !v.height || right > (uint64_t{1} << 32) || bottom > (uint64_t{1} << 32))
        throw std::invalid_argument("histogram requires a nonempty addressable viewport");
    const uint32_t bins = options.bins;
    detail::DeviceBuffer counts(size_t(3) * bins * sizeof(uint64_t)), tallies(9 * sizeof(uint64_t)),
        extrema(6 * sizeof(uint32_t));
    detail::hip(hipMemset(counts.p, 0, size_t(3) * bins * sizeof(uint64_t)), "hipMemset");
    detail::hip(hipMemset(tallies.p, 0, 9 * sizeof(uint64_t)), "hipMemset");
    const uint32_t init[6] = {~0u, 0u, ~0u, 0u, ~0u, 0u};
    detail::hip(hipMemcpy(extrema.p, init, sizeof init, hipMemcpyHostToDevice), "hipMemcpy");
    auto analyse = [&](const float* dev, uint32_t w, uint32_t h) {
        check(rawops_histogram_rgb(dev, w, h, w, bins, options.lower, options.upper,
                                   static_cast<uint64_t*>(counts.p), static_cast<uint64_t*>(tallies.p),
                                   static_cast<uint32_t*>(extrema.p), nullptr));
    };
    if (detail::compilable(output)) {
        detail::validate_rect(detail::request_bounds(source_bounds, request.level), v);
        if (!output.supports_level(request.level)) throw std::invalid_argument("node does not support this render level");
        if (!request.tile_size) throw std::invalid_argument("callback and tile size are required");
        // Bands of whole rows (at least a tile high, up to ~8 M pixels): the
        // histogram does not depend on how the viewport is partitioned.
        const auto chain = detail::compile(output);
        const uint32_t band_h = std::min<uint32_t>(v.height, std::max<uint32_t>(request.tile_size, (8u << 20) / v.width + 1));
        detail::DeviceBuffer dev(size_t(v.width) * band_h * 3 * sizeof(float));
        detail::DeviceBuffer flags(sizeof(uint32_t));
        detail::hip(hipMemset(flags.p, 0, sizeof(uint32_t)), "hipMemset");
        for (uint64_t y = v.y; y < bottom; y += band_h) {
            check_cancelled();
            const Rect band{v.x, uint32_t(y), v.width, uint32_t(std::min<uint64_t>(band_h, bottom - y))};
            detail::run(*chain, band, request.level, static_cast<float*>(dev.p), static_cast<uint32_t*>(flags.p),
                        nullptr);
            check(rawops_check_flags(static_cast<uint32_t*>(flags.p), nullptr));
            analyse(static_cast<const float*>(dev.p), band.width, band.height);
        }
    } else {
        // A foreign node renders through its own code; each tile must be the
        // one the renderer asked for before its storage is read.
        uint64_t next_x = v.x, next_y = v.y;
        // The largest tile the loop can hand out is min(tile_size, viewport)
        // on each side (64-bit product: no wrap for any tile_size).
        const uint64_t tile_w = std::min<uint64_t>(std::max<uint32_t>(1, request.tile_size), v.width);
        const uint64_t tile_h = std::min<uint64_t>(std::max<uint32_t>(1, request.tile_size), v.height);
        detail::DeviceBuffer dev(size_t(tile_w * tile_h * 3 * sizeof(float)));
        Renderer{}.render_tiles(output, source_bounds, request, [&](const Tile& tile) {
            check_cancelled();
            const Rect expected{uint32_t(next_x), uint32_t(next_y),
                                uint32_t(std::min<uint64_t>(request.tile_size, right - next_x)),
                                uint32_t(std::min<uint64_t>(request.tile_size, bottom - next_y))};
            const uint64_t pixels = uint64_t(expected.width) * expected.height;
            if (tile.bounds.x != expected.x || tile.bounds.y != expected.y || tile.bounds.width != expected.width ||
                tile.bounds.height != expected.height || tile.rgb.size() % 3 != 0 || tile.rgb.size() / 3 != pixels ||
                tile.descriptor != result.descriptor)
                throw std::invalid_argument("histogram received inconsistent tile bounds, storage or descriptor");
            detail::hip(hipMemcpy(dev.p, tile.rgb.data(), tile.rgb.size() * sizeof(float), hipMemcpyHostToDevice),
                        "hipMemcpy");
            analyse(static_cast<const float*>(dev.p), expected.width, expected.height);
            next_x += expected.width;
            if (next_x == right) { next_x = v.x; next_y += expected.height; }
            check_cancelled();
        }, cancellation);
    }
    std::vector<uint64_t> c(size_t(3) * bins), t(9);
    uint32_t e[6];
    detail::hip(hipMemcpy(c.data(), counts.p, c.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
    detail::hip(hipMemcpy(t.data(), tallies.p, t.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
    detail::hip(hipMemcpy(e, extrema.p, sizeof e, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int k = 0; k < 3; ++k) {
        auto& ch = result.channels[k];
        ch.counts.assign(c.begin() + size_t(k) * bins, c.begin() + size_t(k + 1) * bins);
        ch.underflow = t[3 * k];
        ch.overflow = t[3 * k + 1];
        ch.nonfinite = t[3 * k + 2];
        if (e[2 * k] != ~0u) {
            ch.minimum = from_order_key(e[2 * k]);
            ch.maximum = from_order_key(e[2 * k + 1]);
        }
    }
    result.pixel_count = uint64_t(v.width) * v.height;
    check_cancelled();
    return result;
}

RgbHistogram histogram_rgb(const ImageGraph& graph, const RenderRequest& request, const HistogramOptions& options,
                           const CancellationToken* cancellation) {
    return histogram_rgb(graph.output(), graph.source_bounds(), request, options, cancellation);
}

}  // namespace rawengine
