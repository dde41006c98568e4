/*
 * CPU oracle for the tiled RAW render path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference RawEngine arithmetic (tylermcm/librawops)
 * used to check the HIP kernels.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path never
 * does.  Every function cites the reference file:line it restates.
 *
 * Pinning status of each function (what independently checks the restatement):
 *   ro_demosaic_bilinear  core_tests KATs + compiled reference (oracle/_ref)
 *   ro_demosaic_menon     93-record frozen golden file menon_native_base_v1.bin
 *                         (bit-exact) + compiled reference on large frames
 *   ro_scale3/ro_scale    closed form (IEEE float multiply) + compiled reference
 *   ro_mat3               closed form + compiled reference
 *   ro_reduce_mip         independent FP64 average (brute force in tests)
 *   ro_tone_curve         closed form + compiled reference
 *   ro_clip / ro_srgb_encode  sRGB KATs (0.04045 -> 0.003130805, ...) + reference
 *   ro_curves_map         monotone-cubic truth values of curves_levels_tests.cpp
 *                         + compiled reference
 *   ro_guided_filter      exact rational impulse KAT [3/22,2/11,5/11,2/11,3/22]
 *                         + brute-force rational kernel + compiled reference (r<=8)
 *   ro_division_mismatches  exhaustive by construction (every operand pair);
 *                         a planted uncorrected quotient fails it
 *   ro_srgb_u8            parity unpinned (no reference implementation exists);
 *                         closed form round(clamp(v,0,1)*255) only.
 */
#ifndef RAWOPS_ORACLE_H
#define RAWOPS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t x, y, width, height; } ro_rect;

typedef struct {
    uint32_t width, height, row_stride;   /* row_stride in samples, already normalised */
    uint32_t pattern;                     /* 0 RGGB, 1 BGGR, 2 GRBG, 3 GBRG */
    uint32_t phase_x, phase_y;
    ro_rect area;                         /* active area, sensor coordinates */
    uint16_t black[4], white[4];          /* pattern-relative 2x2 site order */
} ro_bayer;

/* RawEngine.cpp:299-344 (bilinear branch of RawUnpackNode::render). */
void ro_demosaic_bilinear(const ro_bayer* m, const uint16_t* samples, ro_rect r, float* out);
/* MenonDemosaic.cpp:89-218.  Returns 0 on success, -1 when an axis is a
 * singleton (the reference routes those to bilinear, RawEngine.cpp:303). */
int ro_demosaic_menon(const ro_bayer* m, const uint16_t* samples, ro_rect r, float* out);
/* Menon, row bands rendered in parallel with OpenMP (timing baseline: the
 * "concurrent band requests" configuration of the reference). */
int ro_demosaic_menon_bands(const ro_bayer* m, const uint16_t* samples, ro_rect r,
                            float* out, int threads);

/* FP64 test entry: Menon on double normalised samples (area-relative plane),
 * optionally with frozen classifier decisions (area-relative, 1 =
 * horizontal) and recording the decisions taken; FP64 output.  The taken
 * branch is linear in the samples, which finite differences rely on. */
int ro_menon_f64(const ro_bayer* m, const double* raw, ro_rect r, const uint8_t* hz_in, uint8_t* hz_out,
                 double* out);

/* WhiteBalanceNode::render RawEngine.cpp:417-431. */
void ro_scale3(float* rgb, size_t pixels, const float gains[3]);
/* ExposureNode::render_level RawEngine.cpp:488-500. */
void ro_scale(float* values, size_t count, float multiplier);
/* CameraToWorkingNode::render RawEngine.cpp:532-548 (also the working-space
 * and working-to-sRGB matrices, RawEngine.cpp:636-703). */
void ro_mat3(float* rgb, size_t pixels, const float m[9]);
/* CameraToWorkingNode::render_level RawEngine.cpp:580-601: native is the
 * mapped native rectangle (nw x nh), out is ow x oh at 2^mip reduction. */
void ro_reduce_mip(const float* native, uint32_t nw, uint32_t nh, uint32_t mip,
                   uint32_t ow, uint32_t oh, float* out);
/* ToneCurveNode::render_level RawEngine.cpp:738-758. */
void ro_tone_curve(float* values, size_t count, float shoulder, float inverse_gamma);
/* OutputClipNode::render RawEngine.cpp:1223-1233; returns -1 on non-finite. */
int ro_clip(float* values, size_t count);
/* SrgbEncodeNode::render_level RawEngine.cpp:1254-1277; returns -1 on non-finite. */
int ro_srgb_encode(float* values, size_t count);
/* uint8 quantisation of encoded sRGB (no reference counterpart). */
void ro_srgb_u8(const float* encoded, size_t count, uint8_t* out);

/* Sample normalisation without a divide: for integers n in [0, 65535] and d
 * in [d_lo, d_hi] (1 <= d <= 65535) counts the pairs where
 * q0 + fma(-q0, d, n) * y (fused), q0 = RN(n * y), y = RN(1 / d), differs
 * from the IEEE float quotient n / d that RawEngine.cpp:89-97 takes.  The
 * recipe is odd in n, so n >= 0 covers the normalised range |n| <= 65535.
 * planted != 0 drops the correction (the count must then be non-zero). */
long long ro_division_mismatches(uint32_t d_lo, uint32_t d_hi, int planted);

/* std::exp2 on a float, as ExposureNode computes its multiplier on the host
 * (RawEngine.cpp:465-466). */
float ro_exp2f(float stops);

/* ExtendedCurvesNode ToneOps.cpp:795-844.  knots: x/y pairs, n each (4 curves:
 * master, R, G, B), offsets[c] .. offsets[c+1] index into xs/ys.  cubic selects
 * ShapePreservingCubic.  Returns -1 on a non-finite / out-of-range result. */
int ro_curves(float* rgb, size_t pixels, const double* xs, const double* ys,
              const uint32_t offsets[5], int cubic);

/* working_y_guided_filter_rgb SpatialOps.cpp:168-287 (ordered centred FP64).
 * input covers `needed` = output expanded by 2r clipped to image.  wr/wb are the
 * working-space luminance weights.  Returns 0, or -1 on the reference's
 * overflow/invalid conditions. */
int ro_guided_filter(const float* input, ro_rect needed, ro_rect output, ro_rect image,
                     uint32_t radius, double epsilon, double wr, double wb, float* out);
/* FP64 I/O variant (no float32 rounding of input or output) for finite differences. */
int ro_guided_filter_f64(const double* input, ro_rect needed, ro_rect output, ro_rect image,
                         uint32_t radius, double epsilon, double wr, double wb, double* out);
/* Master curve in FP64 for finite differences (ToneOps.cpp:820-831). */
void ro_curve_map_f64(const double* xs, const double* ys, uint32_t n, int cubic, const double* in, double* out,
                      size_t count);
/* Band-parallel variant of the above for timing (threads OpenMP workers). */
int ro_guided_filter_bands(const float* input, ro_rect needed, ro_rect output, ro_rect image,
                           uint32_t radius, double epsilon, double wr, double wb, float* out,
                           int threads);

/* local_statistics_rgb SpatialOps.cpp:79-121: `in` covers `needed` (output
 * grown by radius, clipped to image); per output pixel and channel the finite
 * count, mean and population variance (NaN moments for an empty window).
 * Row moments by Welford's update in increasing x, merged by Chan's pairwise
 * rule in increasing y. */
void ro_local_statistics(const float* in, ro_rect needed, ro_rect output, ro_rect image, uint32_t radius,
                         double* mean, double* variance, uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif
