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
// LANE > 0 (curves of at most LANE knots): each thread accumulates into its
// own LDS column, so values of one segment in a wavefront (common: the knots
// are few) do not serialise on shared atomics; the columns are summed at the
// end.  LANE == 0: shared LDS atomics.
template <int LANE>
__global__ void __launch_bounds__(kPW) curve_bwd(const float* xv, const float* gy, size_t n, const DevCurve* c,
                                                 uint32_t cubic, float* gx, double* gys) {
    constexpr int KT = LANE > 0 ? LANE : kMaxKnots;          // knots the tables hold
    __shared__ double acc[LANE > 0 ? LANE * kPW : kMaxKnots];
    __shared__ double tab[6 * KT];   // the knot tables (+ 1 / segment width), staged: lookups are data-dependent
    __shared__ double part[LANE > 0 ? (kPW / 64) * LANE : 1];
    for (int k = threadIdx.x; k < (LANE > 0 ? LANE * kPW : kMaxKnots); k += kPW) acc[k] = 0.0;
    const uint32_t nk = as_const(*c).n;
    for (uint32_t k = threadIdx.x; k < nk; k += kPW) {
        tab[k] = c->x[k];
        tab[nk + k] = c->y[k];
        tab[2 * nk + k] = c->slope[k];
        tab[3 * nk + k] = c->c0[k];
        tab[4 * nk + k] = c->c1[k];
        tab[5 * nk + k] = k + 1 < nk ? 1.0 / (c->x[k + 1] - c->x[k]) : 0.0;
    }
    __syncthreads();
    auto add = [&](uint32_t k, double w) {
        if constexpr (LANE > 0) acc[k * kPW + threadIdx.x] += w;
        else atomicAdd(&acc[k], w);
    };
    const LdsCurve l{nk, 0u, tab, tab + nk, tab + 2 * nk, tab + 3 * nk, tab + 4 * nk};
    const LdsCurve* cl = &l;
    const double* inv_h = tab + 5 * nk;   // 1 / (x[k+1] - x[k]): gradients need no quotient rounding
    for (size_t e = blockIdx.x * (size_t)kPW + threadIdx.x; e < n; e += (size_t)gridDim.x * kPW) {
        const double v = xv[e], g = gy[e];
        uint32_t lo = 0, hi = nk;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (v < cl->x[mid]) hi = mid; else lo = mid + 1;
        }
        const uint32_t left = lo == 0 ? 0u : lo - 1u;
        const uint32_t i = left < nk - 2u ? left : nk - 2u;
        const double h = cl->x[i + 1] - cl->x[i], ih = inv_h[i];
        if (v == cl->x[left]) {            // exact knot: value y_left
            add(left, g);
            double slope = cl->slope[left < nk - 1 ? left : nk - 2];   // linear: right-hand secant
            if (cubic) {                                                // cubic: the knot tangent
                slope = left < nk - 1 ? 3.0 * (cl->c0[left] - cl->y[left]) / (cl->x[left + 1] - cl->x[left])
                                      : 3.0 * (cl->y[left] - cl->c1[left - 1]) / (cl->x[left] - cl->x[left - 1]);
            }
            gx[e] = (float)(g * slope);
            continue;
        }
        if (!cubic || v < cl->x[0] || v > cl->x[nk - 1]) {
            // y = y_left + (v - x_left) * s_i,  s_i = (y_i+1 - y_i) / h
            const double dv = v - cl->x[left];
            add(left, g);
            add(i + 1, g * dv * ih);
            add(i, -g * dv * ih);
            gx[e] = (float)(g * cl->slope[i]);
            continue;
        }
        const double t = (v - cl->x[i]) * ih, u = 1.0 - t;
        const double b0 = u * u * u, b1 = 3 * u * u * t, b2 = 3 * u * t * t, b3 = t * t * t;
        const double P0 = cl->y[i], P1 = cl->c0[i], P2 = cl->c1[i], P3 = cl->y[i + 1];
        gx[e] = (float)(g * 3.0 * (u * u * (P1 - P0) + 2 * u * t * (P2 - P1) + t * t * (P3 - P2)) * ih);
        add(i, g * (b0 + b1));
        add(i + 1, g * (b2 + b3));
        // Tangent T_j -> secant slopes (minmod picks one of them, or none).
        auto tangent = [&](uint32_t j, double w) {   // w = dL/dT_j
            if (j == 0) { add(1, w * inv_h[0]); add(0, -w * inv_h[0]); return; }
            if (j == nk - 1) { add(nk - 1, w * inv_h[nk - 2]); add(nk - 2, -w * inv_h[nk - 2]); return; }
            const double a = cl->slope[j - 1], b = cl->slope[j];
            if (a == 0 || b == 0 || signbit(a) != signbit(b)) return;
            const uint32_t k = fabs(a) <= fabs(b) ? j - 1 : j;
            add(k + 1, w * inv_h[k]);
            add(k, -w * inv_h[k]);
        };
        tangent(i, g * b1 * h * (1.0 / 3.0));
        tangent(i + 1, -g * b2 * h * (1.0 / 3.0));
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
