// Per-element forward/adjoint kernels of the differentiable chain: sRGB
// encode (SrgbEncodeNode, RawEngine.cpp:1266-1274), exposure
// (ExposureNode, RawEngine.cpp:494-498), the guided-filter detail boost
// y = x + s (x - GF(x)), and the master tone curve of ExtendedCurvesNode
// (ToneOps.cpp:820-831) with gradients for its knot ordinates.
#include "device_common.hpp"
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kPW = 256;

__device__ __forceinline__ void block_add(double v, double* lds_part, double* out) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kPW / 64; ++w) t += lds_part[w];
        atomicAdd(out, t);
    }
}

__global__ void srgb_fwd(const float* x, float* y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kPW + threadIdx.x; i < n; i += (size_t)gridDim.x * kPW)
        y[i] = srgb_encode(x[i]);
}

// d/dx of clamp01 then 12.92 x | 1.055 x^(1/2.4) - 0.055.
__global__ void srgb_bwd(const float* x, const float* gy, float* gx, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kPW + threadIdx.x; i < n; i += (size_t)gridDim.x * kPW) {
        const float v = x[i];
        float d = 0.0f;
        if (v > 0.0f && v < 1.0f) d = v <= 0.0031308f ? 12.92f : 1.055f / 2.4f * powf(v, 1.0f / 2.4f - 1.0f);
        gx[i] = gy[i] * d;
    }
}

__global__ void exposure_bwd(const float* x, const float* gy, float m, float* gx, double* g_stops, size_t n) {
    __shared__ double part[kPW / 64];
    double acc = 0;
    for (size_t i = blockIdx.x * (size_t)kPW + threadIdx.x; i < n; i += (size_t)gridDim.x * kPW) {
        gx[i] = gy[i] * m;
        acc += (double)gy[i] * (double)x[i] * (double)m;
    }
    block_add(acc * 0.69314718055994530942, part, g_stops);   // d(2^s)/ds = ln2 * 2^s
}

__global__ void boost_fwd(const float* x, const float* base, float s, float* y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kPW + threadIdx.x; i < n; i += (size_t)gridDim.x * kPW)
        y[i] = x[i] + s * (x[i] - base[i]);
}

__global__ void boost_bwd(const float* x, const float* base, const float* gy, float s, float* gx, float* gbase,
                          double* gs, size_t n) {
    __shared__ double part[kPW / 64];
    double acc = 0;
    for (size_t i = blockIdx.x * (size_t)kPW + threadIdx.x; i < n; i += (size_t)gridDim.x * kPW) {
        const float g = gy[i];
        gx[i] = g * (1.0f + s);
        gbase[i] = -s * g;
        acc += (double)g * ((double)x[i] - (double)base[i]);
    }
    block_add(acc, part, gs);
}

// Master-curve adjoint.  The curve at v is one of n + 1 pieces, chosen by the
// number k of knots <= v: the linear extension left of the first knot
// (k = 0), segment k - 1 (1 <= k < n: the cubic with Bernstein control points
// y_i, y_i + h T_i / 3, y_i+1 - h T_i+1 / 3, y_i+1 of a cubic curve, T the
// minmod tangents, each the secant it follows or zero; or the secant of a
// linear curve), the linear extension right of the last knot (k = n).  On
// piece k, with t = (v - x_k') / h_k, the weight of each knot ordinate in the
// value is a polynomial of degree <= 3 in t whose coefficients depend only on
// the knots, and only the four knots k - 2 .. k + 1 have weights: they are
// tabulated per piece once per workgroup, so an element costs one table
// lookup, four Horner evaluations and four adds into its lane's accumulator
// column, with no branches.  dL/dv = g f'(v) from the piece's power-basis
// coefficients (f = c0 + c1 u + c2 u^2 + c3 u^3, u = v - x_k').  At a knot
// (t = 0) the weights and the derivative are those of the piece starting
// there: the value y_j exactly and the knot tangent (cubic) or right-hand
// secant (linear), as the forward's exact-knot rule gives (ToneOps.cpp:780-792).
// The piece is found by a search over float thresholds (for a float v,
// v >= x_k exactly when v >= the smallest float >= x_k).  LANE: each thread
// adds into its own LDS column (curves of at most 16 knots), so values of one
// piece in a wavefront do not serialise on shared atomics; else shared LDS
// atomics.  EXPO: the curve follows an exposure v = m x (the forward's FP32
// product); the kernel reads x, writes dL/dx = m (g f'(v)) and adds
// ln2 sum(g f'(v) m x) into *gst, the values exposure_bwd gives from the
// curve adjoint's output, without that output's round trip through memory.
constexpr int kRec = 6;    // per piece: x', d1 = c1, d2 = 2 c2, d3 = 3 c3, 1 / h, pad
constexpr int kWts = 16;   // per piece: knot k - 2 + q, power p of t: [q][p]

// Minmod tangent source of knot j (ToneOps.cpp:795-816): -1 the secant j - 1,
// 0 the secant j, 1 none (zero tangent).
__device__ __forceinline__ int tangent_source(Const<DevCurve>& c, uint32_t j) {
    const uint32_t n = c.n;
    if (j == 0) return 0;
    if (j == n - 1) return -1;
    const double a = c.slope[j - 1], b = c.slope[j];
    return (a == 0 || b == 0 || signbit(a) != signbit(b)) ? 1 : (fabs(a) <= fabs(b) ? -1 : 0);
}

template <bool LANE, bool EXPO>
__global__ void __launch_bounds__(kPW) curve_bwd(const float* xv, const float* gy, size_t n, const DevCurve* c,
                                                 uint32_t cubic, float* gx, double* gys, float m, double* gst) {
    constexpr int KMAX = LANE ? 16 : kMaxKnots;                 // knots the tables hold
    constexpr int COLS = KMAX + 4;                              // knot columns -2 .. KMAX + 1
    __shared__ __attribute__((aligned(16))) double rec[kRec * (KMAX + 1)];
    __shared__ __attribute__((aligned(16))) double wts[kWts * (KMAX + 1)];
    __shared__ __attribute__((aligned(16))) float xf[KMAX < 16 ? 16 : KMAX];
    __shared__ double acc[LANE ? COLS * kPW : COLS];
    __shared__ double part[LANE ? (kPW / 64) * KMAX : 1];
    Const<DevCurve>& s = as_const(*c);
    const uint32_t nk = s.n;
    uint32_t p2 = 4;
    while (p2 < nk) p2 <<= 1;
    const float nan = __builtin_nanf("");
    const float t4 = nk > 4 ? float_at_least(s.x[4]) : nan, t8 = nk > 8 ? float_at_least(s.x[8]) : nan,
                t12 = nk > 12 ? float_at_least(s.x[12]) : nan;
    for (uint32_t k = threadIdx.x; k < p2; k += kPW) xf[k] = k < nk ? float_at_least(s.x[k]) : nan;
    for (uint32_t k = threadIdx.x; k < (LANE ? COLS * kPW : COLS); k += kPW) acc[k] = 0.0;
    for (uint32_t k = threadIdx.x; k <= nk; k += kPW) {
        double* r = rec + kRec * k;
        double* w = wts + kWts * k;   // w[4 q + p]: knot k - 2 + q, t^p
        for (int q = 0; q < kWts; ++q) w[q] = 0.0;
        double x, d1, d2 = 0.0, d3 = 0.0, h;
        if (k == 0) {                  // y0 + t (y1 - y0)
            x = s.x[0]; d1 = s.slope[0]; h = s.x[1] - s.x[0];
            w[4 * 2 + 0] = 1.0; w[4 * 2 + 1] = -1.0;
            w[4 * 3 + 1] = 1.0;
        } else if (k == nk) {          // y_n-1 + t (y_n-1 - y_n-2)
            x = s.x[nk - 1]; d1 = s.slope[nk - 2]; h = s.x[nk - 1] - s.x[nk - 2];
            w[4 * 0 + 1] = -1.0;
            w[4 * 1 + 0] = 1.0; w[4 * 1 + 1] = 1.0;
        } else {
            const uint32_t i = k - 1;  // knots i - 1 .. i + 2 are q = 0 .. 3
            x = s.x[i]; h = s.x[i + 1] - s.x[i];
            if (cubic) {
                const double y0 = s.y[i], y1 = s.y[i + 1], p = s.c0[i], q = s.c1[i], inv = 1.0 / h;
                d1 = (3.0 * (p - y0)) * inv;
                d2 = 2.0 * (((3.0 * ((y0 - 2.0 * p) + q)) * inv) * inv);
                d3 = 3.0 * (((((y1 - y0) + 3.0 * (p - q)) * inv) * inv) * inv);
                // b0 + b1 = 1 - 3t^2 + 2t^3 on y_i, b2 + b3 = 3t^2 - 2t^3 on y_i+1
                w[4 * 1 + 0] = 1.0; w[4 * 1 + 2] = -3.0; w[4 * 1 + 3] = 2.0;
                w[4 * 2 + 2] = 3.0; w[4 * 2 + 3] = -2.0;
                // (h / 3) b1 T_i, b1 / 3 = t - 2t^2 + t^3; T_i a secant: d/dy = +-1 / its width
                const double B1[4] = {0.0, 1.0, -2.0, 1.0};
                const int si = tangent_source(s, i);
                if (si == 0) {
                    for (int e = 0; e < 4; ++e) { w[4 * 2 + e] += B1[e]; w[4 * 1 + e] -= B1[e]; }
                } else if (si == -1) {
                    const double r = h / (s.x[i] - s.x[i - 1]);
                    for (int e = 0; e < 4; ++e) { w[4 * 1 + e] += r * B1[e]; w[4 * 0 + e] -= r * B1[e]; }
                }
                // -(h / 3) b2 T_i+1, b2 / 3 = t^2 - t^3
                const double B2[4] = {0.0, 0.0, -1.0, 1.0};
                const int si1 = tangent_source(s, i + 1);
                if (si1 == -1) {
                    for (int e = 0; e < 4; ++e) { w[4 * 2 + e] += B2[e]; w[4 * 1 + e] -= B2[e]; }
                } else if (si1 == 0) {
                    const double r = h / (s.x[i + 2] - s.x[i + 1]);
                    for (int e = 0; e < 4; ++e) { w[4 * 3 + e] += r * B2[e]; w[4 * 2 + e] -= r * B2[e]; }
                }
            } else {                   // y_i + t (y_i+1 - y_i)
                d1 = s.slope[i];
                w[4 * 1 + 0] = 1.0; w[4 * 1 + 1] = -1.0;
                w[4 * 2 + 1] = 1.0;
            }
        }
        r[0] = x; r[1] = d1; r[2] = d2; r[3] = d3; r[4] = 1.0 / h; r[5] = 0.0;
    }
    __syncthreads();
    auto piece_of = [&](float v) -> uint32_t {
        if (p2 <= 16) {
            const uint32_t g = (v >= t4 ? 1u : 0u) + (v >= t8 ? 1u : 0u) + (v >= t12 ? 1u : 0u);
            const float4 q = reinterpret_cast<const float4*>(xf)[g];
            return 4u * g + (v >= q.x ? 1u : 0u) + (v >= q.y ? 1u : 0u) + (v >= q.z ? 1u : 0u) + (v >= q.w ? 1u : 0u);
        }
        uint32_t cnt = 0;
        for (uint32_t st = p2 >> 1; st > 0; st >>= 1) cnt += v >= xf[cnt + st - 1] ? st : 0u;
        return cnt + (v >= xf[cnt] ? 1u : 0u);
    };
    double sacc = 0.0;   // EXPO: sum of g f'(v) m x
    auto one = [&](size_t e, float xe, double g) {
        const float vf = EXPO ? xe * m : xe;
        const uint32_t k = piece_of(vf);
        const double2* r = reinterpret_cast<const double2*>(rec + kRec * k);
        const double2 a = r[0], b = r[1], hh = r[2];
        const double du = (double)vf - a.x;
        const float gc = (float)(g * __builtin_fma(__builtin_fma(b.y, du, b.x), du, a.y));
        if constexpr (EXPO) {
            gx[e] = gc * m;
            sacc += (double)gc * (double)xe * (double)m;
        } else {
            gx[e] = gc;
        }
        const double t = du * hh.x;
        const double2* w = reinterpret_cast<const double2*>(wts + kWts * k);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double2 lo = w[2 * q], hi = w[2 * q + 1];
            const double wq = g * __builtin_fma(__builtin_fma(__builtin_fma(hi.y, t, hi.x), t, lo.y), t, lo.x);
            const int col = (int)k + q;   // column of knot k - 2 + q
            if constexpr (LANE) acc[col * kPW + threadIdx.x] += wq;
            else if (wq != 0.0) atomicAdd(&acc[col], wq);
        }
    };
    // Four elements per thread per round, their loads issued together.
    const size_t stride = (size_t)gridDim.x * kPW;
    for (size_t e0 = blockIdx.x * (size_t)kPW + threadIdx.x; e0 < n; e0 += 4 * stride) {
        float vv[4], gg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t e = min(e0 + u * stride, n - 1);
            vv[u] = xv[e];
            gg[u] = gy[e];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * stride < n) one(e0 + u * stride, vv[u], (double)gg[u]);
    }
    if constexpr (LANE) {
        // sum the columns of knots 0 .. nk - 1: per knot a wavefront reduction, then the waves
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (uint32_t k = 0; k < nk; ++k) {
            double v = acc[(k + 2) * kPW + threadIdx.x];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) part[wave * KMAX + k] = v;
        }
        __syncthreads();
        if (threadIdx.x < nk) {
            double t = 0;
            for (int w = 0; w < kPW / 64; ++w) t += part[w * KMAX + threadIdx.x];
            atomicAdd(gys + threadIdx.x, t);
        }
    } else {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nk; k += kPW) atomicAdd(gys + k, acc[k + 2]);
    }
    if constexpr (EXPO) {
        __shared__ double spart[kPW / 64];
        block_add(sacc * 0.69314718055994530942, spart, gst);   // d(2^s)/ds = ln2 * 2^s
    }
}

unsigned blocks_for(size_t n) { return (unsigned)std::min<size_t>((n + kPW - 1) / kPW, 8192); }
// Kernels that end in one FP64 atomic per block: at least 8 elements per
// thread, so small inputs do not serialise thousands of atomics on the one
// result (large inputs keep the full streaming grid).
unsigned reduce_blocks_for(size_t n) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 8 * kPW - 1) / (8 * kPW), 8192));
}

}  // namespace

hipError_t launch_srgb(const float* x, float* y, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(srgb_fwd, dim3(blocks_for(n)), dim3(kPW), 0, s, x, y, n);
    return hipGetLastError();
}
hipError_t launch_srgb_backward(const float* x, const float* gy, float* gx, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(srgb_bwd, dim3(blocks_for(n)), dim3(kPW), 0, s, x, gy, gx, n);
    return hipGetLastError();
}
hipError_t launch_exposure_backward(const float* x, const float* gy, float m, float* gx, double* g_stops, size_t n,
                                    hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(exposure_bwd, dim3(reduce_blocks_for(n)), dim3(kPW), 0, s, x, gy, m, gx, g_stops, n);
    return hipGetLastError();
}
hipError_t launch_boost(const float* x, const float* base, float strength, float* y, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(boost_fwd, dim3(blocks_for(n)), dim3(kPW), 0, s, x, base, strength, y, n);
    return hipGetLastError();
}
hipError_t launch_boost_backward(const float* x, const float* base, const float* gy, float strength, float* gx,
                                 float* gbase, double* gs, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(boost_bwd, dim3(reduce_blocks_for(n)), dim3(kPW), 0, s, x, base, gy, strength, gx, gbase, gs, n);
    return hipGetLastError();
}
hipError_t launch_curve_backward(const float* x, const float* gy, size_t n, const DevCurve* curve, uint32_t knots,
                                 uint32_t cubic, float* gx, double* gys, hipStream_t s, float m, double* g_stops) {
    if (!n) return hipSuccess;
    const bool lane = knots <= 16;   // per-thread accumulator columns
    const dim3 grid(lane ? reduce_blocks_for(n) : blocks_for(n));
    if (g_stops) {
        if (lane) hipLaunchKernelGGL((curve_bwd<true, true>), grid, dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys, m, g_stops);
        else hipLaunchKernelGGL((curve_bwd<false, true>), grid, dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys, m, g_stops);
    } else {
        if (lane) hipLaunchKernelGGL((curve_bwd<true, false>), grid, dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys, 1.0f, nullptr);
        else hipLaunchKernelGGL((curve_bwd<false, false>), grid, dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys, 1.0f, nullptr);
    }
    return hipGetLastError();
}

}  // namespace rawops
