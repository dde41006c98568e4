// Device-side building blocks shared by the render, guided-filter and adjoint
// kernels: sensor description, the fused per-pixel op program and its
// interpreter.  Compiled for gfx950 only, with FP contraction disabled so every
// float/double operation rounds exactly where the reference's does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rawops {

constexpr int kMaxOps = 12;
constexpr int kMaxKnots = 256;

enum OpKind : uint32_t {
    OP_GAINS = 1, OP_SCALE = 2, OP_MATRIX = 3, OP_TONE = 4,
    OP_CLIP = 5, OP_SRGB_ENCODE = 6, OP_CURVES = 7
};

// Error bits OR-ed into the caller's device flag word.
enum : uint32_t { FLAG_DOMAIN = 1u, FLAG_INVALID = 2u, FLAG_OVERFLOW = 4u };

// One prepared curve (ExtendedCurvesNode ctor, ToneOps.cpp:795-816).
struct DevCurve {
    uint32_t n, identity;
    double x[kMaxKnots], y[kMaxKnots];
    double slope[kMaxKnots], c0[kMaxKnots], c1[kMaxKnots];
};
struct DevCurves {
    uint32_t cubic;
    DevCurve c[4];  // master, R, G, B
};

struct DevOp {
    uint32_t kind;
    float v[9];
    const DevCurves* curves;
    uint32_t knots;          // curve ops: knot count of the master curve (host-side launch choices)
};

struct OpList {
    uint32_t n;
    DevOp op[kMaxOps];
};

// Plan-owned tables (op lists, curves) are written by the host before a
// launch and only read by kernels: reading them through the constant address
// space lets uniform reads go through the scalar cache (no vector-memory wait).
template <class T>
using Const = __attribute__((address_space(4))) const T;
template <class T>
__device__ __forceinline__ Const<T>& as_const(const T& v) {
    return *(Const<T>*)&v;
}

// Bayer sensor as the kernels see it (RawMetadata after normalisation).
struct SensorArgs {
    const uint16_t* samples;
    uint32_t stride;          // samples per row
    int ax, ay, aw, ah;       // active area, sensor coordinates
    uint32_t colors;          // colour (0 R, 1 G, 2 B) of site s in bits 2s..2s+1
    uint32_t phase_x, phase_y;
    uint32_t ng_parity;       // bit b: sensor x parity of the non-green sites in rows of parity b
    float black[4], denom[4]; // per site: black level, white - black
    float rcp[4];             // per site: 1 / (white - black), correctly rounded to float
};

__device__ __forceinline__ uint32_t site_of(const SensorArgs& s, int x, int y) {
    return ((((uint32_t)y & 1u) + s.phase_y) & 1u) * 2u + ((((uint32_t)x & 1u) + s.phase_x) & 1u);
}

__device__ __forceinline__ int color_of(const SensorArgs& s, int x, int y) {
    return (int)((s.colors >> (2u * site_of(s, x, y))) & 3u);
}

__device__ __forceinline__ bool in_area(const SensorArgs& s, int x, int y) {
    return x >= s.ax && y >= s.ay && x < s.ax + s.aw && y < s.ay + s.ah;
}

// Per-site value by select (no indexed access, so the sensor copy stays in
// registers).
__device__ __forceinline__ float site_value(const float (&v)[4], uint32_t site) {
    const bool hi = (site & 2u) != 0, odd = (site & 1u) != 0;
    return hi ? (odd ? v[3] : v[2]) : (odd ? v[1] : v[0]);
}

// (code - black) / (white - black) rounded exactly as the IEEE float
// division of RawEngine.cpp:89-97 rounds it: n = code - black and
// d = white - black are integers with |n| <= 65535 and 1 <= d <= 65535, and
// for every such pair the corrected product q0 + (n - q0 d) y, with
// y = RN(1/d), q0 = RN(n y) and both corrections fused, equals RN(n / d)
// (checked exhaustively over every such pair by the oracle's
// ro_division_mismatches).
__device__ __forceinline__ float normalize_exact(float code, float black, float rcp, float denom) {
    const float n = code - black;
    const float q0 = n * rcp;
    const float r = __builtin_fmaf(-q0, denom, n);
    return __builtin_fmaf(r, rcp, q0);
}

__device__ __forceinline__ float normalize_code(const SensorArgs& s, uint32_t code, uint32_t site) {
    return normalize_exact((float)code, site_value(s.black, site), site_value(s.rcp, site), site_value(s.denom, site));
}

// Whole-sample reflection into [origin, origin + size), size >= 2, for
// coordinates a few samples outside (MenonDemosaic.cpp:60-66; repeated
// reflection equals its periodic fold).
__device__ __forceinline__ int reflect(int v, int origin, int size) {
    const int last = origin + size - 1;
    while (v < origin || v > last) v = v < origin ? 2 * origin - v : 2 * last - v;
    return v;
}

__device__ __forceinline__ float clamp01(float v) {
    // std::clamp(v, 0.0f, 1.0f): keeps -0.0f, as the reference does.
    return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);
}

__device__ __forceinline__ bool finite(float v) { return __builtin_isfinite(v); }

__device__ __forceinline__ float tone(float linear, float shoulder, float inverse_gamma) {
    const float magnitude = fabsf(linear);
    const float compressed = magnitude / (magnitude + shoulder);
    const float mapped = inverse_gamma == 1.0f ? compressed : powf(compressed, inverse_gamma);
    return copysignf(mapped, linear);
}

// SrgbEncodeNode's transfer function (RawEngine.cpp:1264-1275).  The power
// segment's x^(1/2.4), x in (0.0031308, 1], is exp2(log2(x) / 2.4) from the
// hardware base-2 logarithm and exponential (v_log_f32, v_exp_f32: three
// instructions instead of a full-range powf): its deviation from the
// correctly rounded power is a few 1e-7, inside the 1e-5 bar of the encoded
// output (checked over every eighth float of the segment on the GPU).
__device__ __forceinline__ float srgb_encode(float v) {
    const float linear = clamp01(v);
    const float power = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(linear) * (1.0f / 2.4f));
    return linear <= 0.0031308f ? 12.92f * linear : 1.055f * power - 0.055f;
}

// curve_cubic_blend, ToneOps.cpp:788-792.
__device__ __forceinline__ double cubic_blend(double a, double b, double t) {
    if (a == b || t == 0) return a;
    if (t == 1) return b;
    return t <= .5 ? a + t * (b - a) : b + (1 - t) * (a - b);
}

// A curve's knot tables staged in LDS (same fields as DevCurve, n knots).
struct LdsCurve {
    uint32_t n, identity;
    const double *x, *y, *slope, *c0, *c1;
};

// ExtendedCurvesNode::map, ToneOps.cpp:820-831, over a DevCurve in the
// constant address space or an LdsCurve.
template <class C>
__device__ inline double curve_map_in(const C& c, uint32_t cubic, double v) {
    if (c.identity) return v;
    // upper_bound on x (v is finite, x strictly increasing).  Short curves:
    // count the knots <= v, so every knot load has the same address on all
    // lanes and none waits on another; long curves: binary search.
    uint32_t lo = 0;
    if (c.n <= 32) {
#pragma unroll 8
        for (uint32_t k = 0; k < c.n; ++k) lo += v >= c.x[k] ? 1u : 0u;
    } else {
        uint32_t hi = c.n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (v < c.x[mid]) hi = mid; else lo = mid + 1;
        }
    }
    const uint32_t left = lo == 0 ? 0u : lo - 1u;
    const uint32_t i = left < c.n - 2u ? left : c.n - 2u;
    if (v == c.x[left]) return c.y[left];
    if (!cubic || v < c.x[0] || v > c.x[c.n - 1])
        return c.slope[i] == 0 ? c.y[left] : c.y[left] + (v - c.x[left]) * c.slope[i];
    const double t = (v - c.x[i]) / (c.x[i + 1] - c.x[i]);
    const double a = cubic_blend(c.y[i], c.c0[i], t), b = cubic_blend(c.c0[i], c.c1[i], t),
                 d = cubic_blend(c.c1[i], c.y[i + 1], t);
    return cubic_blend(cubic_blend(a, b, t), cubic_blend(b, d, t), t);
}

__device__ inline double curve_map(const DevCurve& curve, uint32_t cubic, double v) {
    return curve_map_in(as_const(curve), cubic, v);
}

struct CurveOut {
    float v;
    uint32_t err;
};

// Out of line: the op program unrolls over a thread's pixels, and the curve
// lookup is long.
static __device__ __noinline__ CurveOut curves_channel_v(const DevCurves* cv, int channel, float value) {
    if (cv->c[0].identity && cv->c[channel + 1].identity) return {value, 0u};
    if (!finite(value)) return {value, FLAG_INVALID};
    const double master = curve_map(cv->c[0], cv->cubic, (double)value);
    if (!__builtin_isfinite(master)) return {value, FLAG_INVALID};
    const double mapped = curve_map(cv->c[channel + 1], cv->cubic, master);
    if (!__builtin_isfinite(mapped) || fabs(mapped) > 3.4028234663852886e38) return {value, FLAG_INVALID};
    return {(float)mapped, 0u};
}

// Inline form for the per-pixel op program.
__device__ inline float curves_channel(const DevCurves* cv, int channel, float value, uint32_t& err) {
    if (cv->c[0].identity && cv->c[channel + 1].identity) return value;
    if (!finite(value)) { err |= FLAG_INVALID; return value; }
    const double master = curve_map(cv->c[0], cv->cubic, (double)value);
    if (!__builtin_isfinite(master)) { err |= FLAG_INVALID; return value; }
    const double mapped = curve_map(cv->c[channel + 1], cv->cubic, master);
    if (!__builtin_isfinite(mapped) || fabs(mapped) > 3.4028234663852886e38) { err |= FLAG_INVALID; return value; }
    return (float)mapped;
}

// A curve staged in LDS for the filter passes: per knot count k (the number
// of knots <= v, 0..n) one record {x_k', c0, c1, c2, c3} with which the
// curve at v is c0 + u (c1 + u (c2 + u c3)), u = v - x_k' (three fused
// multiply-adds, no selects):
//   k = 0       left of the first knot, the linear extension
//               (x_0, y_0, slope_0, 0, 0);
//   1 <= k < n  segment i = k - 1 from x_i: the cubic (cubic curves: control
//               points y_i, c0_i, c1_i, y_i+1; power-basis coefficients in
//               t = u / h, q0 = y_i, q1 = 3 (c0 - y_i), q2 = 3 (y_i - 2 c0 + c1),
//               q3 = y_i+1 - y_i + 3 (c0 - c1), divided by 1, h, h^2, h^3) or
//               the secant (y_i, slope_i, 0, 0);
//   k = n       at or right of the last knot, the linear extension
//               (x_n-1, y_n-1, slope_n-2, 0, 0).
// u = 0 (v at a knot) returns c0 = y exactly, and a zero slope returns y:
// the exact-knot and flat-segment rules of curve_map_in.  The records bake in
// the knot selection, the linear extensions and the cubic / linear choice
// of curve_map_in; cubic values agree with its de Casteljau blends
// (ToneOps.cpp:788-792) to FP64 rounding (the stored results are float32).
//
// For the knot search of float inputs the curve also keeps, per knot, the
// smallest float >= x_k (for a float v, v >= x_k exactly when v >= that
// float), padded with NaN to a power of two p2 >= max(n, 4), so that the
// number of knots <= v is a search over floats: for p2 <= 16 two levels
// (thresholds 4, 8 and 12 held in registers pick a group of four, whose
// thresholds are one 16-byte LDS read), a binary search otherwise.
struct LdsPoly {
    uint32_t n, identity;
    const double* x;     // the n knot abscissae
    const double* rec;   // n + 1 records of kPolyRec doubles, 16-byte aligned
    const float* xf;     // p2 float thresholds, 16-byte aligned
    uint32_t p2;
    float t4, t8, t12;   // xf[4], xf[8], xf[12] (NaN at k >= n)
};

constexpr int kPolyRec = 6;   // x, c0, c1, c2, c3, pad

// LDS doubles of one staged curve of at most cap knots, cap a power of two
// (even, so that consecutive curves stay 16-byte aligned).
constexpr int poly_table_doubles(int cap) { return (kPolyRec * (cap + 1) + cap + 1 + cap / 2 + 1) & ~1; }

// The smallest float >= x (x finite; +-FLT_MAX overflow to +-inf is handled
// by the step toward +inf).
__device__ __forceinline__ float float_at_least(double x) {
    float f = (float)x;
    if ((double)f < x) {
        const uint32_t b = __float_as_uint(f);
        f = __uint_as_float((b & 0x80000000u) ? b - 1u : b + 1u);
    }
    return f;
}

// Stage curve c (2 <= n <= cap knots) at base (poly_table_doubles(cap)
// doubles, 16-byte aligned) for the cubic or linear form; the whole
// workgroup works, the caller synchronises.
__device__ __forceinline__ LdsPoly stage_poly(const DevCurve* c, uint32_t cubic, double* base) {
    Const<DevCurve>& s = as_const(*c);
    const uint32_t n = s.n;
    uint32_t p2 = 4;
    while (p2 < n) p2 <<= 1;
    double* x = base + kPolyRec * (n + 1);
    float* xf = reinterpret_cast<float*>(base + ((kPolyRec * (n + 1) + n + 1) & ~1u));
    const float nan = __builtin_nanf("");
    LdsPoly l{n, s.identity, x, base, xf, p2,
              n > 4 ? float_at_least(s.x[4]) : nan, n > 8 ? float_at_least(s.x[8]) : nan,
              n > 12 ? float_at_least(s.x[12]) : nan};
    if (l.identity) return l;
    for (uint32_t k = threadIdx.x; k < p2; k += blockDim.x) xf[k] = k < n ? float_at_least(s.x[k]) : nan;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) x[k] = s.x[k];
    for (uint32_t k = threadIdx.x; k <= n; k += blockDim.x) {
        double* r = base + kPolyRec * k;
        double c2 = 0.0, c3 = 0.0;
        if (k == 0) {
            r[0] = s.x[0]; r[1] = s.y[0]; r[2] = s.slope[0];
        } else if (k == n) {
            r[0] = s.x[n - 1]; r[1] = s.y[n - 1]; r[2] = s.slope[n - 2];
        } else {
            const uint32_t i = k - 1;
            r[0] = s.x[i];
            r[1] = s.y[i];
            if (cubic) {
                const double y0 = s.y[i], y1 = s.y[i + 1], p = s.c0[i], q = s.c1[i];
                const double inv = 1.0 / (s.x[i + 1] - s.x[i]);
                r[2] = (3.0 * (p - y0)) * inv;
                c2 = ((3.0 * ((y0 - 2.0 * p) + q)) * inv) * inv;
                c3 = ((((y1 - y0) + 3.0 * (p - q)) * inv) * inv) * inv;
            } else {
                r[2] = s.slope[i];
            }
        }
        r[3] = c2;
        r[4] = c3;
        r[5] = 0.0;
    }
    return l;
}

// The curve at v given cnt = the number of knots <= v (no branches).
__device__ __forceinline__ double poly_eval(const LdsPoly& c, double v, uint32_t cnt) {
    const double2* r = reinterpret_cast<const double2*>(c.rec + kPolyRec * cnt);
    const double2 a = r[0], b = r[1], d = r[2];
    const double u = v - a.x;
    return __builtin_fma(__builtin_fma(__builtin_fma(d.x, u, b.y), u, b.x), u, a.y);
}

// Number of knots <= v for a float v: a binary search over the float
// thresholds (NaN padding and a NaN v compare false).
__device__ __forceinline__ uint32_t knot_count_f(const LdsPoly& c, float v) {
    if (c.p2 <= 16) {
        const uint32_t g = (v >= c.t4 ? 1u : 0u) + (v >= c.t8 ? 1u : 0u) + (v >= c.t12 ? 1u : 0u);
        const float4 q = reinterpret_cast<const float4*>(c.xf)[g];
        return 4u * g + (v >= q.x ? 1u : 0u) + (v >= q.y ? 1u : 0u) + (v >= q.z ? 1u : 0u) + (v >= q.w ? 1u : 0u);
    }
    uint32_t cnt = 0;
    for (uint32_t s = c.p2 >> 1; s > 0; s >>= 1) cnt += v >= c.xf[cnt + s - 1] ? s : 0u;
    return cnt + (v >= c.xf[cnt] ? 1u : 0u);
}

// ExtendedCurvesNode::map (ToneOps.cpp:820-831) for the three channels of a
// pixel over LDS-staged curves (master curves[0], channel curves[1..3]):
// the checks and flags of curves_channel.
__device__ inline void curves3_lds(const LdsPoly* curves, float o[3], uint32_t& err) {
    double v[3], m[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) v[ch] = (double)o[ch];
    if (!curves[0].identity) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) m[ch] = poly_eval(curves[0], v[ch], knot_count_f(curves[0], o[ch]));
    } else {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) m[ch] = v[ch];
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const LdsPoly& cc = curves[ch + 1];
        if (curves[0].identity && cc.identity) continue;
        double mapped = m[ch];
        if (!cc.identity) {
            uint32_t k = 0;
            for (uint32_t j = 0; j < cc.n; ++j) k += m[ch] >= cc.x[j] ? 1u : 0u;
            mapped = poly_eval(cc, m[ch], k);
        }
        const bool bad = !finite(o[ch]) || !__builtin_isfinite(m[ch]) || !__builtin_isfinite(mapped) ||
                         fabs(mapped) > 3.4028234663852886e38;
        if (bad) err |= FLAG_INVALID;
        o[ch] = bad ? o[ch] : (float)mapped;
    }
}

// The fused op program: each op rounds exactly like its reference node.
__device__ __forceinline__ void apply_ops(const OpList& list, uint32_t begin, uint32_t end, float& r, float& g, float& b,
                                          uint32_t& err) {
    Const<OpList>& L = as_const(list);
    for (uint32_t i = begin; i < end; ++i) {
        Const<DevOp>& o = L.op[i];
        switch (o.kind) {
        case OP_GAINS: r *= o.v[0]; g *= o.v[1]; b *= o.v[2]; break;
        case OP_SCALE: r *= o.v[0]; g *= o.v[0]; b *= o.v[0]; break;
        case OP_MATRIX: {
            const float R = o.v[0] * r + o.v[1] * g + o.v[2] * b;
            const float G = o.v[3] * r + o.v[4] * g + o.v[5] * b;
            const float B = o.v[6] * r + o.v[7] * g + o.v[8] * b;
            r = R; g = G; b = B;
            break;
        }
        case OP_TONE:
            r = tone(r, o.v[0], o.v[1]); g = tone(g, o.v[0], o.v[1]); b = tone(b, o.v[0], o.v[1]);
            break;
        case OP_CLIP:
            if (!finite(r) || !finite(g) || !finite(b)) err |= FLAG_DOMAIN;
            r = clamp01(r); g = clamp01(g); b = clamp01(b);
            break;
        case OP_SRGB_ENCODE:
            if (!finite(r) || !finite(g) || !finite(b)) err |= FLAG_DOMAIN;
            r = srgb_encode(r); g = srgb_encode(g); b = srgb_encode(b);
            break;
        case OP_CURVES:
            r = curves_channel(o.curves, 0, r, err);
            g = curves_channel(o.curves, 1, g, err);
            b = curves_channel(o.curves, 2, b, err);
            break;
        default: break;
        }
    }
}

__device__ __forceinline__ void apply_ops(const OpList& L, float& r, float& g, float& b, uint32_t& err) {
    apply_ops(L, 0, as_const(L).n, r, g, b, err);
}

// The op program over a thread's pixels, op by op: one (uniform) dispatch
// per op for all NPX pixels.  err[p] collects pixel p's error bits.
template <int NPX>
__device__ __forceinline__ void apply_ops_px(const OpList& list, uint32_t begin, uint32_t end, float (&px)[NPX][3],
                                             uint32_t (&err)[NPX]) {
    Const<OpList>& L = as_const(list);
    for (uint32_t i = begin; i < end; ++i) {
        Const<DevOp>& o = L.op[i];
        switch (o.kind) {
        case OP_GAINS: {
            const float a = o.v[0], b = o.v[1], c = o.v[2];
#pragma unroll
            for (int p = 0; p < NPX; ++p) { px[p][0] *= a; px[p][1] *= b; px[p][2] *= c; }
            break;
        }
        case OP_SCALE: {
            const float a = o.v[0];
#pragma unroll
            for (int p = 0; p < NPX; ++p) { px[p][0] *= a; px[p][1] *= a; px[p][2] *= a; }
            break;
        }
        case OP_MATRIX: {
            float m[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) m[k] = o.v[k];
#pragma unroll
            for (int p = 0; p < NPX; ++p) {
                const float r = px[p][0], g = px[p][1], b = px[p][2];
                px[p][0] = m[0] * r + m[1] * g + m[2] * b;
                px[p][1] = m[3] * r + m[4] * g + m[5] * b;
                px[p][2] = m[6] * r + m[7] * g + m[8] * b;
            }
            break;
        }
        case OP_TONE: {
            const float sh = o.v[0], ig = o.v[1];
#pragma unroll
            for (int p = 0; p < NPX; ++p)
#pragma unroll
                for (int c = 0; c < 3; ++c) px[p][c] = tone(px[p][c], sh, ig);
            break;
        }
        case OP_CLIP:
#pragma unroll
            for (int p = 0; p < NPX; ++p) {
                if (!finite(px[p][0]) || !finite(px[p][1]) || !finite(px[p][2])) err[p] |= FLAG_DOMAIN;
#pragma unroll
                for (int c = 0; c < 3; ++c) px[p][c] = clamp01(px[p][c]);
            }
            break;
        case OP_SRGB_ENCODE:
#pragma unroll
            for (int p = 0; p < NPX; ++p) {
                if (!finite(px[p][0]) || !finite(px[p][1]) || !finite(px[p][2])) err[p] |= FLAG_DOMAIN;
#pragma unroll
                for (int c = 0; c < 3; ++c) px[p][c] = srgb_encode(px[p][c]);
            }
            break;
        case OP_CURVES:
#pragma unroll
            for (int p = 0; p < NPX; ++p)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const CurveOut r = curves_channel_v(o.curves, c, px[p][c]);
                    px[p][c] = r.v;
                    err[p] |= r.err;
                }
            break;
        default: break;
        }
    }
}

__device__ __forceinline__ uint8_t quantize_u8(float v) {
    return (uint8_t)(clamp01(v) * 255.0f + 0.5f);
}

}  // namespace rawops
