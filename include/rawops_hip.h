/*
 * rawops_hip.h -- C ABI of the MI355X (gfx950) tiled RAW render path.
 *
 * This is the boundary the C++ drop-in layer (include/rawengine/RawEngine.hpp,
 * rawengine::ImageGraph / Renderer) and the Python/PyTorch module call.  It
 * replaces, for the render hot path, the node kernels the reference runs
 * under Renderer::render_tiles (RawEngine.cpp:1390-1412):
 *   - RawUnpackNode::render (bilinear RawEngine.cpp:299-344, Menon
 *     MenonDemosaic.cpp:177-218),
 *   - WhiteBalanceNode/ExposureNode/CameraToWorkingNode/WorkingSpaceConvert/
 *     WorkingToSrgb/ToneCurve/OutputClip/SrgbEncode (RawEngine.cpp:408-758,
 *     1215-1277) and ExtendedCurvesNode (ToneOps.cpp:795-844) as a fused
 *     per-pixel op program,
 *   - the mip-1/2 footprint reduction of CameraToWorkingNode::render_level and
 *     RasterSourceNode::render_level (RawEngine.cpp:363-406, 580-601),
 *   - working_y_guided_filter_rgb (SpatialOps.cpp:168-287).
 * Plain pointers and sizes only.  `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Device pointers are HBM addresses; host pointers
 * are only taken where the name says so.  No exceptions cross this boundary:
 * every call returns a rawops_status and rawops_last_error() holds the message
 * (per thread).  Status codes map one to one onto the reference's exception
 * types (RawEngine.hpp:646-649 and the std:: types it throws).
 */
#ifndef RAWOPS_HIP_H
#define RAWOPS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RAWOPS_OK = 0,
    RAWOPS_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    RAWOPS_OUT_OF_RANGE = 2,     /* std::out_of_range (viewport outside source) */
    RAWOPS_DOMAIN_ERROR = 3,     /* std::domain_error (non-finite at an output boundary) */
    RAWOPS_LENGTH_ERROR = 4,     /* std::length_error */
    RAWOPS_OVERFLOW_ERROR = 5,   /* std::overflow_error (guided filter) */
    RAWOPS_CANCELLED = 6,        /* rawengine::RenderCancelled */
    RAWOPS_LOGIC_ERROR = 7,      /* std::logic_error */
    RAWOPS_DEVICE_ERROR = 8      /* HIP runtime failure */
} rawops_status;

typedef struct { uint32_t x, y, width, height; } rawops_rect;

/* RawMetadata, RawEngine.hpp:57-65 (same fields, same normalisation rules:
 * stride 0 = width, all-zero active area = full sensor). */
typedef struct {
    uint32_t width, height, row_stride_samples;
    uint32_t pattern;                 /* 0 RGGB, 1 BGGR, 2 GRBG, 3 GBRG */
    uint32_t cfa_phase_x, cfa_phase_y;
    rawops_rect active_area;
    uint16_t black_levels[4], white_levels[4];
} rawops_raw_metadata;

enum { RAWOPS_DEMOSAIC_BILINEAR = 0, RAWOPS_DEMOSAIC_MENON = 1 };
enum { RAWOPS_PROPHOTO_D50 = 0, RAWOPS_REC2020_D65 = 1 };
enum { RAWOPS_QUALITY_PREVIEW = 0, RAWOPS_QUALITY_FINAL = 1 };

/* Per-pixel operations of the fused op program. */
enum {
    RAWOPS_OP_GAINS = 1,        /* rgb *= v[0..2]                WhiteBalanceNode */
    RAWOPS_OP_SCALE = 2,        /* rgb *= v[0]                   ExposureNode */
    RAWOPS_OP_MATRIX = 3,       /* (v0*r + v1*g) + v2*b ...      CameraToWorking etc. */
    RAWOPS_OP_TONE = 4,         /* v0 shoulder, v1 inverse gamma ToneCurveNode */
    RAWOPS_OP_CLIP = 5,         /* non-finite -> domain error    OutputClipNode */
    RAWOPS_OP_SRGB_ENCODE = 6,  /* non-finite -> domain error    SrgbEncodeNode */
    RAWOPS_OP_CURVES = 7        /* knots below                   ExtendedCurvesNode */
};

typedef struct {
    uint32_t kind;
    float v[9];
    /* RAWOPS_OP_CURVES: master, R, G, B knot lists concatenated (host memory,
     * copied by rawops_plan_create); knot_offsets[c]..knot_offsets[c+1]. */
    const double* knots_x;
    const double* knots_y;
    uint32_t knot_offsets[5];
    uint32_t cubic;
} rawops_op_desc;

/* A compiled chain: demosaic, ops at native resolution (up to and including
 * the reduction anchor), ops after the mip reduction.  For mip 0 both lists
 * run back to back.  `mip_anchor` is nonzero when the chain has a reduction
 * anchor (CameraToWorkingNode with native bounds, RawEngine.cpp:550-601);
 * reduced levels are rejected without one. */
typedef struct {
    uint32_t demosaic;
    uint32_t mip_anchor;
    uint32_t n_native, n_reduced;
    const rawops_op_desc* native_ops;
    const rawops_op_desc* reduced_ops;
} rawops_plan_desc;

typedef struct rawops_source rawops_source;
typedef struct rawops_plan rawops_plan;

const char* rawops_last_error(void);
const char* rawops_version(void);

/* ---- colour math on the host (RawEngine.cpp:116-219, 502-530, 603-672) ---- */
/* CameraToWorkingNode ctor: float32 camera->working matrix, row-major. */
rawops_status rawops_camera_to_working_matrix(const double camera_to_xyz_d50[9],
                                              uint32_t working_space, float out[9]);
/* WorkingToSrgbNode ctor: working -> linear sRGB. */
rawops_status rawops_working_to_srgb_matrix(uint32_t working_space, float out[9]);
/* WorkingSpaceConvertNode ctor: source -> target working space. */
rawops_status rawops_working_to_working_matrix(uint32_t source, uint32_t target, float out[9]);
/* FP64 XYZ(D50) -> working-space matrix (RawEngine.cpp:183-196), row-major. */
rawops_status rawops_xyz_d50_to_working_matrix(uint32_t working_space, double out[9]);
/* ExposureNode multiplier, std::exp2(stops) in float with the reference's
 * stop range check (RawEngine.cpp:465-468). */
rawops_status rawops_exposure_multiplier(float stops, float* out);

/* ---- sources (device-resident) ---- */
/* RawImage ctor validation (RawEngine.cpp:223-250) + upload of host samples. */
rawops_status rawops_source_create_raw(const rawops_raw_metadata* metadata,
                                       const uint16_t* host_samples, size_t sample_count,
                                       rawops_source** out);
/* Same validation; wraps caller-owned device samples without copying. */
rawops_status rawops_source_wrap_raw_device(const rawops_raw_metadata* metadata,
                                            const uint16_t* device_samples, size_t sample_count,
                                            rawops_source** out);
/* RasterImage (RawEngine.hpp:96-119): interleaved float32 RGB, stride in pixels. */
rawops_status rawops_source_create_raster(uint32_t width, uint32_t height, uint32_t row_stride_pixels,
                                          uint32_t working_space, const float* host_pixels,
                                          rawops_source** out);
rawops_status rawops_source_wrap_raster_device(uint32_t width, uint32_t height,
                                               uint32_t row_stride_pixels, uint32_t working_space,
                                               const float* device_pixels, rawops_source** out);
/* Normalised metadata of a RAW source (stride and active area filled in). */
rawops_status rawops_source_raw_metadata(const rawops_source* source, rawops_raw_metadata* out);
void rawops_source_destroy(rawops_source* source);

/* ---- plans ---- */
rawops_status rawops_plan_create(const rawops_plan_desc* desc, rawops_plan** out);
void rawops_plan_destroy(rawops_plan* plan);

/* Source bounds of a request level (request_bounds, RawEngine.cpp:31-43). */
rawops_status rawops_request_bounds(rawops_rect source_bounds, uint32_t mip, uint32_t quality,
                                    rawops_rect* out);

/* Render `viewport` (coordinates at level `mip`) of a RAW source through
 * `plan` into device memory: dst_rgb (float32 interleaved RGB, row stride
 * dst_stride_pixels, may be NULL) and/or dst_u8 (uint8 sRGB quantisation of
 * the final values, round(clamp(v,0,1)*255), row stride dst_stride_pixels*3
 * bytes, may be NULL).  Asynchronous on `stream`.  device_flags (may be NULL)
 * is a device uint32 the kernels OR error bits into (bit 0: non-finite at an
 * output boundary); rawops_check_flags maps it to a status. */
rawops_status rawops_render(const rawops_plan* plan, const rawops_source* source,
                            rawops_rect viewport, uint32_t mip, uint32_t quality,
                            float* dst_rgb, uint8_t* dst_u8, uint32_t dst_stride_pixels,
                            uint32_t* device_flags, void* stream);
/* Render path of Menon mip-0 renders (no reference counterpart: a tuning
 * switch, process-wide).  path 0: automatic (the streaming kernel for
 * viewports large enough to fill the device, about 2.75 MP and up, the tile
 * kernel below); 1: the streaming kernel wherever the active area is at
 * least 16 x 16; 2: the tile kernel only; 3: as 1, with the Menon stages of
 * linear chains (white balance, exposure, 3x3) computed in FP32 instead of
 * the reference's FP64: an approximation whose classifier decisions and
 * outputs differ from the reference's (bench.py reports its flip rate and
 * largest deviation).  band_rows: rows each streaming wavefront renders
 * (0 = 16, else a multiple of 4 in 4..4096).  Paths 0, 1 and 2 render the
 * same bits. */
rawops_status rawops_set_menon_path(uint32_t path, uint32_t band_rows);

/* Reads a device_flags word (synchronising `stream`) and resets it to zero. */
rawops_status rawops_check_flags(uint32_t* device_flags, void* stream);

/* Synchronous host render with the reference's tile loop semantics
 * (Renderer::render_tiles RawEngine.cpp:1390-1412): the viewport is rendered on
 * the device in one launch, then `callback(user, tile_rect, rgb)` receives each
 * tile_size x tile_size tile in row-major order from host memory; a nonzero
 * *cancel (polled before the render and before each tile) returns
 * RAWOPS_CANCELLED.  A nonzero callback return stops the loop with that value
 * passed through as RAWOPS_CANCELLED.  The callback's pointer is valid only
 * during the call. */
typedef int (*rawops_tile_callback)(void* user, rawops_rect tile, const float* rgb);
rawops_status rawops_render_tiles_host(const rawops_plan* plan, const rawops_source* source,
                                       rawops_rect viewport, uint32_t mip, uint32_t quality,
                                       uint32_t tile_size, rawops_tile_callback callback,
                                       void* user, const volatile int* cancel);

/* ---- batched heterogeneous render (the TileScheduler caller's job,
 * TileScheduler.cpp:19-246, as one device launch) ---- */
typedef struct {
    const rawops_plan* plan;
    const rawops_source* source;       /* RAW source */
    rawops_rect viewport;              /* level coordinates */
    uint32_t mip, quality;
    uint32_t linear_ops;               /* reduced ops applied before dst_rgb; the rest precede dst_u8 */
    float* dst_rgb;                    /* device, may be NULL */
    uint8_t* dst_u8;                   /* device, may be NULL */
    uint32_t dst_stride_pixels;        /* 0 = viewport width */
} rawops_batch_request;

typedef struct rawops_batch rawops_batch;

/* Validates every request (as rawops_render does), builds the (request, tile)
 * work list, picks the tile shape (shape_index < rawops_batch_shape_count():
 * 64x32, 64x64, 128x32, one workgroup per (request, tile) item; or
 * UINT32_MAX to autotune: every shape is timed on this batch and the fastest
 * kept) and captures reset + tile launch into a hipGraph (captured on a
 * private stream, replayed on the caller's).  Requests are served smallest
 * first. */
rawops_status rawops_batch_create(const rawops_batch_request* requests, size_t count, uint32_t shape_index,
                                  void* stream, rawops_batch** out);
/* Replays the captured graph (asynchronous on `stream`). */
rawops_status rawops_batch_run(rawops_batch* batch, void* stream);
/* After the stream has completed a run: per-request completion latency (ms)
 * from the start of the launch, in request order; and non-finite status. */
rawops_status rawops_batch_latencies(rawops_batch* batch, double* latency_ms);
rawops_status rawops_batch_check(rawops_batch* batch, void* stream);
/* Chosen tile shape and its autotune time (ms per batch, 0 if not tuned). */
rawops_status rawops_batch_shape(const rawops_batch* batch, uint32_t* tile_width, uint32_t* tile_height,
                                 double* tuned_ms);
uint32_t rawops_batch_shape_count(void);
void rawops_batch_destroy(rawops_batch* batch);

/* ---- adjoints (no reference counterpart: the reference has no derivatives) ---- */
/* Reverse mode of the render chain out = M * (m * (g (.) demosaic(raw)))
 * (WhiteBalanceNode -> ExposureNode -> CameraToWorkingNode after the demosaic,
 * RawEngine.cpp:408-548) at mip 0.  Given dL/dout over `viewport` (device,
 * float32 RGB, row stride g_out_stride_pixels), writes dL/d(normalised raw
 * sample) into g_raw (device float,
 * active_width x active_height, row-major) and ADDS dL/d(gains[3],
 * multiplier, matrix[9]) into g_params (device double[13]).  Only photosites
 * within 4 of the viewport (the adjoint's reach) are written, so a caller
 * that needs the whole plane zero-fills it first; g_raw may be NULL when only
 * the parameter gradients are wanted.  Bilinear: exact
 * transpose (remosaic scatter).  Menon: transpose of the branch the forward
 * classifier takes (decisions recomputed, not differentiated). */
rawops_status rawops_render_backward(const rawops_source* source, uint32_t demosaic, rawops_rect viewport,
                                     const float gains[3], float multiplier, const float matrix[9],
                                     const float* g_out, uint32_t g_out_stride_pixels, float* g_raw,
                                     double* g_params, void* stream);

/* Menon classifier decisions recorded by the forward for the adjoint (no
 * reference counterpart).  A decision record of a viewport (vx, vy, vw, vh)
 * holds one bit per non-green photosite, 1 = horizontal, 32 words per
 * 64 x 32 tile of the viewport: with X = x - vx, Y = y - vy the site at
 * sensor (x, y) is bit (X / 2) % 32 of the 32-bit word
 * ((Y / 32) * ((vw + 63) / 64) + X / 64) * 32 + Y % 32 (device memory,
 * 4-byte aligned).  The size in bytes that holds the record of any viewport: */
rawops_status rawops_menon_decisions_size(const rawops_source* source, size_t* bytes);
/* rawops_render of a Menon plan at mip 0 (FP32 output only) that also writes
 * the decision record of `viewport` into `decisions` (device, every word of
 * the viewport's record is written).  Output bits equal rawops_render's. */
rawops_status rawops_render_record(const rawops_plan* plan, const rawops_source* source, rawops_rect viewport,
                                   float* dst_rgb, uint32_t dst_stride_pixels, uint32_t* device_flags,
                                   uint8_t* decisions, void* stream);
/* rawops_render_backward (Menon) with the decisions a rawops_render_record of
 * the viewport `recorded` wrote: tiles whose adjoint reach lies inside
 * `recorded` and away from the area edge read them instead of recomputing the
 * classifier (the rest recompute as rawops_render_backward does).  Same
 * outputs as rawops_render_backward. */
rawops_status rawops_render_backward_recorded(const rawops_source* source, rawops_rect viewport,
                                              const float gains[3], float multiplier, const float matrix[9],
                                              const float* g_out, uint32_t g_out_stride_pixels, float* g_raw,
                                              double* g_params, const uint8_t* decisions, rawops_rect recorded,
                                              void* stream);

/* Reverse mode of the working-Y guided filter (radius 1..32, level 0) of a
 * raster source: given dL/dq over `viewport` (device float32 RGB, row stride
 * g_out_stride_pixels), writes dL/dp for every raster pixel into g_in (device
 * float32, width x height x 3) and ADDS dL/d(eps) into *g_eps (device
 * double).  The per-centre box statistics are recomputed from the input in
 * FP64; workspace is device scratch of rawops_guided_backward_workspace_size()
 * bytes (NULL: allocated stream-ordered by the call). */
rawops_status rawops_guided_backward_workspace_size(uint32_t raster_width, uint32_t raster_height, size_t* bytes);
rawops_status rawops_guided_backward(const rawops_source* raster, uint32_t radius, double epsilon,
                                     rawops_rect viewport, const float* g_out, uint32_t g_out_stride_pixels,
                                     float* g_in, double* g_eps, void* workspace, size_t workspace_bytes,
                                     void* stream);

/* Element-wise forward/adjoint kernels (device pointers, n floats). */
/* sRGB encode of clamp01(x) (SrgbEncodeNode, RawEngine.cpp:1266-1274) and its derivative. */
rawops_status rawops_srgb_encode(const float* x, float* y, size_t n, void* stream);
rawops_status rawops_srgb_encode_backward(const float* x, const float* g_y, float* g_x, size_t n, void* stream);
/* y = m x with m = 2^stops: g_x = m g_y, *g_stops += ln2 * sum(g_y * m x). */
rawops_status rawops_exposure_backward(const float* x, const float* g_y, float multiplier, float* g_x,
                                       double* g_stops, size_t n, void* stream);
/* Detail boost y = x + s (x - base) and its adjoint (*g_s accumulated). */
rawops_status rawops_detail_boost(const float* x, const float* base, float strength, float* y, size_t n, void* stream);
rawops_status rawops_detail_boost_backward(const float* x, const float* base, const float* g_y, float strength,
                                           float* g_x, float* g_base, double* g_s, size_t n, void* stream);
/* Adjoint of the master curve of a plan whose reduced_ops[0] is RAWOPS_OP_CURVES
 * (ExtendedCurvesNode::map, ToneOps.cpp:820-831): g_x and the knot-ordinate
 * gradients g_knots_y (device double[knot count], accumulated). */
rawops_status rawops_curve_backward(const rawops_plan* curve_plan, const float* x, const float* g_y, size_t n,
                                    float* g_x, double* g_knots_y, void* stream);
/* ExposureNode (y = m x, m from rawops_exposure_multiplier) followed by the
 * master curve: the adjoints of both in one pass over x (the exposure input)
 * and g_y.  g_x, g_knots_y and *g_stops as rawops_curve_backward on m x
 * followed by rawops_exposure_backward give them (g_x bit for bit; the FP64
 * sums up to their order), without the intermediate gradient in memory.  No
 * reference counterpart: derivatives are new. */
rawops_status rawops_exposure_curve_backward(const rawops_plan* curve_plan, const float* x, float multiplier,
                                             const float* g_y, size_t n, float* g_x, double* g_knots_y,
                                             double* g_stops, void* stream);

/* ---- guided filter (SpatialOps.cpp:159-317) + fused point epilogue ---- */
/* Working-Y guided filter of a raster source at level `mip`, output rect
 * `viewport` (level coordinates), radius 0..32 (the reference admits 0..8,
 * SpatialOps.cpp:159-163; larger radii are an extension), eps in
 * [2^-24, 65536], then the plan's `reduced_ops` program per pixel (post may be
 * NULL).  Variance/covariance box moments are accumulated in FP64.
 * `workspace` is device scratch of rawops_guided_workspace_size() bytes for
 * the per-centre coefficients; NULL makes the call allocate it stream-ordered. */
rawops_status rawops_guided_workspace_size(uint32_t raster_width, uint32_t raster_height,
                                           rawops_rect viewport, uint32_t mip, uint32_t radius,
                                           size_t* bytes);
rawops_status rawops_guided_filter(const rawops_source* raster, rawops_rect viewport,
                                   uint32_t mip, uint32_t radius, double epsilon,
                                   const rawops_plan* post, float* dst_rgb,
                                   uint32_t dst_stride_pixels, void* workspace,
                                   size_t workspace_bytes, uint32_t* device_flags,
                                   void* stream);

/* rawops_guided_filter for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The per-centre coefficient
 * workspace lives in a static device array of
 * RAWOPS_GUIDED_STATIC_SCRATCH_BYTES per device (held until the process
 * ends) instead of caller or stream-ordered memory; the call itself
 * allocates nothing.  A request whose workspace
 * (rawops_guided_workspace_size) exceeds that size returns
 * RAWOPS_LENGTH_ERROR before any launch.  Every call on a device shares the
 * one array: calls that overlap on a device (concurrent streams or threads)
 * race and must be serialised by the caller.  Same outputs as
 * rawops_guided_filter, bit for bit. */
#define RAWOPS_GUIDED_STATIC_SCRATCH_BYTES (64u << 20)
rawops_status rawops_guided_filter_static(const rawops_source* raster, rawops_rect viewport,
                                          uint32_t mip, uint32_t radius, double epsilon,
                                          const rawops_plan* post, float* dst_rgb,
                                          uint32_t dst_stride_pixels, uint32_t* device_flags,
                                          void* stream);

/* ---- analysis (ImageAnalysis.cpp:10-70) ---- */
/* Per-channel histogram of a device FP32 RGB image (width x height, row
 * stride stride_pixels, 0 = width), accumulated into device counters so that
 * an image can be analysed in bands: counts[c * bins + k] (uint64), tallies[3c
 * + 0..2] = underflow, overflow, nonfinite (uint64), extrema[2c], [2c + 1] =
 * order keys of the finite minimum and maximum (uint32; initialise them to
 * 0xFFFFFFFF and 0; key(v) = bits(v) ^ 0x80000000 for v >= +0 and ~bits(v)
 * for v <= -0, so keys sort as the floats do).  Bins: `bins` (1..65536) equal
 * bins of [lower, upper], half-open except the last, which includes upper;
 * bin = min(bins - 1, (uint32)((double(v) - lower) / (upper - lower) * bins)).
 * Values below lower / above upper are tallied, not clipped; NaN and
 * infinities count as nonfinite and do not enter the extrema.  Replaces the
 * per-tile loop of histogram_rgb (ImageAnalysis.cpp:47-63). */
rawops_status rawops_histogram_rgb(const float* rgb, uint32_t width, uint32_t height, uint32_t stride_pixels,
                                   uint32_t bins, double lower, double upper, uint64_t* counts,
                                   uint64_t* tallies, uint32_t* extrema, void* stream);

/* Local mean, population variance and finite-sample count per channel
 * (local_statistics_rgb, SpatialOps.cpp:72-121) of a device FP32 RGB tile
 * covering `needed` = `output` grown by `radius` (0..8) and clipped to
 * `image` (level coordinates; row stride in pixels, 0 = needed.width).
 * Windows clip at the image edge; non-finite samples are skipped; windows
 * without finite samples give NaN moments.  FP64 Welford row moments merged
 * in increasing row order, the reference's order: results equal its FP64
 * values bit for bit.  mean / variance (double) and finite_count (uint32) are
 * output.width x output.height x 3, row-major, on the device.  Replaces
 * the per-tile computation of analyze_local_rgb (SpatialOps.cpp:123-146). */
rawops_status rawops_local_statistics(const float* rgb, rawops_rect needed, uint32_t stride_pixels,
                                      rawops_rect output, rawops_rect image, uint32_t radius, double* mean,
                                      double* variance, uint32_t* finite_count, void* stream);


/* A stack of `images` equal-size images (no reference counterpart; for
 * callers that filter many crops of one size, such as the fit over crops):
 * the raster's height splits into `images` images of height H each,
 * consecutive in memory, and each is filtered on its own (level 0, its own
 * borders) in one launch, with the same results as rawops_guided_filter /
 * rawops_guided_backward on each image.  Outputs are consecutive too: image i
 * of dst_rgb / g_out starts i * viewport.height * stride pixels in, image i of
 * g_in i * width * H pixels in.  Workspaces are allocated stream-ordered. */
rawops_status rawops_guided_filter_stack(const rawops_source* raster, uint32_t images, rawops_rect viewport,
                                         uint32_t radius, double epsilon, const rawops_plan* post,
                                         float* dst_rgb, uint32_t dst_stride_pixels, uint32_t* device_flags,
                                         void* stream);
rawops_status rawops_guided_backward_stack(const rawops_source* raster, uint32_t images, uint32_t radius,
                                           double epsilon, rawops_rect viewport, const float* g_out,
                                           uint32_t g_out_stride_pixels, float* g_in, double* g_eps,
                                           void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RAWOPS_HIP_H */
