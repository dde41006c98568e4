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

// Master-curve adjoint.  For the segment the value falls in, the cubic is the
// Bezier B(t) = sum_k b_k(t) P_k with P = (y_i, y_i + h T_i/3, y_i+1 - h T_i+1/3,
// y_i+1); tangents T are the minmod of the adjacent secant slopes.  dL/dy_j
// collects the Bernstein weights through P and T; dL/dx uses B'(t)/h.
// Per element every contribution lands on the knots i-1 .. i+2 (the tangent
// of knot j follows the secant k_j in {j-1, j} that the minmod picks, or
// none, fixed by the curve and tabulated once per block), so the weights of
// those four knots are formed in registers and added with four
// read-modify-writes.  The segment is found by a binary search over float
// thresholds: for a float v, v < x_k exactly when v < the smallest float >=
// x_k.  LANE > 0 (curves of at most LANE knots): each thread accumulates into
// its own LDS column, so values of one segment in a wavefront (common: the
// knots are few) do not serialise on shared atomics; the columns are summed
// at the end.  LANE == 0: shared LDS atomics.
template <int LANE>
__global__ void __launch_bounds__(kPW) curve_bwd(const float* xv, const float* gy, size_t n, const DevCurve* c,
                                                 uint32_t cubic, float* gx, double* gys) {
    constexpr int KT = LANE > 0 ? LANE : kMaxKnots;          // knots the tables hold
    __shared__ double acc[LANE > 0 ? LANE * kPW : kMaxKnots];
    __shared__ double tab[6 * KT];   // the knot tables (+ 1 / segment width), staged: lookups are data-dependent
    __shared__ float xf[KT];         // the float thresholds, +inf beyond the last knot
    __shared__ int src[KT];          // k_j - j (-1 or 0): the secant knot j's tangent follows, 1 for none
    __shared__ double part[LANE > 0 ? (kPW / 64) * LANE : 1];
    for (int k = threadIdx.x; k < (LANE > 0 ? LANE * kPW : kMaxKnots); k += kPW) acc[k] = 0.0;
    const uint32_t nk = as_const(*c).n;
    uint32_t p2 = 1;
    while (p2 < nk) p2 <<= 1;
    for (uint32_t k = threadIdx.x; k < p2; k += kPW) {
        xf[k] = k < nk ? float_at_least(c->x[k]) : __builtin_inff();
        if (k >= nk) continue;
        tab[k] = c->x[k];
        tab[nk + k] = c->y[k];
        tab[2 * nk + k] = c->slope[k];
        tab[3 * nk + k] = c->c0[k];
        tab[4 * nk + k] = c->c1[k];
        tab[5 * nk + k] = k + 1 < nk ? 1.0 / (c->x[k + 1] - c->x[k]) : 0.0;
        int sk;
        if (k == 0) {
            sk = 0;
        } else if (k == nk - 1) {
            sk = -1;
        } else {
            const double a = c->slope[k - 1], b = c->slope[k];
            sk = (a == 0 || b == 0 || signbit(a) != signbit(b)) ? 1 : (fabs(a) <= fabs(b) ? -1 : 0);
        }
        src[k] = sk;
    }
    __syncthreads();
    const double *X = tab, *Y = tab + nk, *SL = tab + 2 * nk, *C0 = tab + 3 * nk, *C1 = tab + 4 * nk;
    const double* inv_h = tab + 5 * nk;   // 1 / (x[k+1] - x[k]): gradients need no quotient rounding
    auto one = [&](size_t e, float vf, double g) {
        const double v = vf;
        // lo = the number of knots x_k with !(v < x_k) (all of them for NaN)
        uint32_t lo = 0;
        for (uint32_t step = p2 >> 1; step > 0; step >>= 1) lo += !(vf < xf[lo + step - 1]) ? step : 0u;
        lo += !(vf < xf[lo]) ? 1u : 0u;
        lo = lo < nk ? lo : nk;
        const uint32_t left = lo == 0 ? 0u : lo - 1u;
        const uint32_t i = left < nk - 2u ? left : nk - 2u;
        const double ih = inv_h[i];
        // dL/dy of the knots i-1, i, i+1, i+2
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
        const double gl = g, gl1 = left == i ? gl : 0.0, gl2 = left == i ? 0.0 : gl;   // dL/dy_left
        float gxe;
        if (v == X[left]) {               // exact knot: value y_left
            w1 = gl1;
            w2 = gl2;
            double slope = SL[left < nk - 1 ? left : nk - 2];   // linear: right-hand secant
            if (cubic) {                                       // cubic: the knot tangent
                slope = left < nk - 1 ? 3.0 * (C0[left] - Y[left]) / (X[left + 1] - X[left])
                                      : 3.0 * (Y[left] - C1[left - 1]) / (X[left] - X[left - 1]);
            }
            gxe = (float)(g * slope);
        } else if (!cubic || v < X[0] || v > X[nk - 1]) {
            // y = y_left + (v - x_left) * s_i,  s_i = (y_i+1 - y_i) / h
            const double dv = v - X[left];
            const double q = g * dv * ih;
            w1 = gl1 - q;
            w2 = gl2 + q;
            gxe = (float)(g * SL[i]);
        } else {
            const double h = X[i + 1] - X[i];
            const double t = (v - X[i]) * ih, u = 1.0 - t;
            const double b0 = u * u * u, b1 = 3 * u * u * t, b2 = 3 * u * t * t, b3 = t * t * t;
            const double P0 = Y[i], P1 = C0[i], P2 = C1[i], P3 = Y[i + 1];
            gxe = (float)(g * 3.0 * (u * u * (P1 - P0) + 2 * u * t * (P2 - P1) + t * t * (P3 - P2)) * ih);
            w1 = g * (b0 + b1);
            w2 = g * (b2 + b3);
            // Tangent T_j -> secant k (minmod picks one of the adjacent secants,
            // or none): dL/dy_k+1 += dT / h_k, dL/dy_k -= dT / h_k.
            const double Ti = g * b1 * h * (1.0 / 3.0), Ti1 = -g * b2 * h * (1.0 / 3.0);
            const int si = src[i], si1 = src[i + 1];
            const double a = si == -1 ? Ti * inv_h[i > 0 ? i - 1 : 0] : 0.0;   // secant i-1
            const double b = si == 0 ? Ti * ih : 0.0;                            // secant i
            const double cc = si1 == -1 ? Ti1 * ih : 0.0;                        // secant i
            const double d = si1 == 0 ? Ti1 * inv_h[i + 1 < nk - 1 ? i + 1 : i] : 0.0;   // secant i+1
            w0 = -a;
            w1 += a - b - cc;
            w2 += b + cc - d;
            w3 = d;
        }
        gx[e] = gxe;
        const double w[4] = {w0, w1, w2, w3};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = (int)i - 1 + q;
            if (j < 0 || j >= (int)nk) continue;
            if constexpr (LANE > 0) acc[j * kPW + threadIdx.x] += w[q];
            else if (w[q] != 0.0) atomicAdd(&acc[j], w[q]);
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
    if constexpr (LANE > 0) {
        // sum the columns: per knot a wavefront reduction, then the waves
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (uint32_t k = 0; k < nk; ++k) {
            double v = acc[k * kPW + threadIdx.x];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) part[wave * LANE + k] = v;
        }
        __syncthreads();
        if (threadIdx.x < nk) {
            double t = 0;
            for (int w = 0; w < kPW / 64; ++w) t += part[w * LANE + threadIdx.x];
            atomicAdd(gys + threadIdx.x, t);
        }
    } else {
        __syncthreads();
        for (int k = threadIdx.x; k < (int)nk; k += kPW) atomicAdd(gys + k, acc[k]);
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
                                 uint32_t cubic,
                                 float* gx, double* gys, hipStream_t s) {
    if (!n) return hipSuccess;
    if (knots <= 16)   // per-thread accumulator columns
        hipLaunchKernelGGL(curve_bwd<16>, dim3(reduce_blocks_for(n)), dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys);
    else
        hipLaunchKernelGGL(curve_bwd<0>, dim3(blocks_for(n)), dim3(kPW), 0, s, x, gy, n, curve, cubic, gx, gys);
    return hipGetLastError();
}

}  // namespace rawops
