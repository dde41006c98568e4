// Per-channel histogram of an FP32 RGB image (histogram_rgb,
// ImageAnalysis.cpp:10-70): counts over `bins` equal bins of [lower, upper]
// (half-open, the last bin closed), underflow / overflow / nonfinite tallies
// and the finite extrema, accumulated into device counters so that an image
// can be analysed band by band.
//
// One pass over the pixels (a band of rows per workgroup).  With up
// to kLdsBins bins a workgroup counts into LDS and adds its nonzero bins to
// the global counters once; wider histograms count with global atomics.  The
// tallies and extrema are reduced per wave with shuffles and added by one
// lane per wave.  Extrema are kept as order keys
// (unsigned integers that sort as the floats do) so that atomicMin/Max apply.
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kHistThreads = 256;
constexpr uint32_t kLdsBins = 2048;
// Pixels whose loads each thread has in flight together.  Four keep the
// kernel within 128 VGPRs (four waves per SIMD, 119 VGPRs, no scratch); eight
// took 221 VGPRs and two waves per SIMD, and forcing those into 128 spilled.
constexpr int kPix = 4;

__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <bool IN_LDS>
__global__ void __launch_bounds__(kHistThreads) __attribute__((amdgpu_waves_per_eu(4))) histogram_rgb(const HistogramLaunch h) {
    __shared__ uint32_t lh[IN_LDS ? 3 * kLdsBins : 1];
    constexpr bool in_lds = IN_LDS;
    const uint32_t nb = 3 * h.bins;
    if (in_lds) {
        for (uint32_t i = threadIdx.x; i < nb; i += kHistThreads) lh[i] = 0;
        __syncthreads();
    }
    uint32_t under[3] = {0, 0, 0}, over[3] = {0, 0, 0}, nonfinite[3] = {0, 0, 0};
    uint32_t kmin[3] = {~0u, ~0u, ~0u}, kmax[3] = {0, 0, 0};
    const double bins = (double)h.bins;
    // The bin is trunc((sample - lower) / span * bins) with a correctly
    // rounded division.  A product by the reciprocal differs from that
    // quotient times bins by at most 5 units in the last place (two more
    // roundings on each side), so where it lies farther than a relative 1e-13
    // from every bin edge both truncate alike; samples nearer an edge take the
    // division.  Reciprocals outside [1e-300, 1e300] always divide.
    const double inv = 1.0 / h.span;
    const bool by_inverse = inv >= 1e-300 && inv <= 1e300;
    // a band of whole rows per workgroup, threads across each row; every
    // thread issues the loads of kPix pixels before it counts any of them
    const uint32_t rows = (h.height + gridDim.x - 1) / gridDim.x;
    const uint32_t y0 = blockIdx.x * rows, y1 = min(h.height, y0 + rows);
    for (uint32_t y = y0; y < y1; ++y) {
        const float* row = h.rgb + (size_t)y * h.stride * 3;
        for (uint32_t x0 = 0; x0 < h.width; x0 += kHistThreads * kPix) {
            float v[kPix][3];
            bool live[kPix];
#pragma unroll
            for (int p = 0; p < kPix; ++p) {
                const uint32_t x = x0 + threadIdx.x + p * kHistThreads;
                live[p] = x < h.width;
                const float* q = row + (size_t)min(x, h.width - 1) * 3;
                v[p][0] = q[0]; v[p][1] = q[1]; v[p][2] = q[2];
            }
            // Every sample's bin from the reciprocal product, without branches
            // (so that the samples' FP64 chains overlap); the division only
            // for samples near an edge, after the batch, when a lane has one.
            uint32_t bin[kPix][3], inc[kPix][3], near = 0;
#pragma unroll
            for (int p = 0; p < kPix; ++p)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float s = v[p][c];
                const bool fin = live[p] && __builtin_isfinite(s);
                nonfinite[c] += live[p] && !fin;
                const uint32_t k = order_key(s);
                kmin[c] = fin ? min(kmin[c], k) : kmin[c];
                kmax[c] = fin ? max(kmax[c], k) : kmax[c];
                const double d = (double)s;
                const bool lo = fin && d < h.lower, hi = fin && d > h.upper;
                under[c] += lo;
                over[c] += hi;
                const bool counted = fin && !lo && !hi;
                const double t1 = (d - h.lower) * inv * bins;
                const double e = rint(t1);
                const bool edge = !by_inverse || (e >= 1.0 && e < bins && fabs(t1 - e) <= 1e-13 * t1);
                near |= (uint32_t)(counted && edge) << (p * 3 + c);
                bin[p][c] = min(h.bins - 1, counted ? (uint32_t)t1 : 0u);
                inc[p][c] = counted;
            }
            if (near) {
#pragma unroll
                for (int p = 0; p < kPix; ++p)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (near >> (p * 3 + c) & 1u)     // Divide before multiplying (no overflow for extreme ranges).
                        bin[p][c] = min(h.bins - 1, (uint32_t)(((double)v[p][c] - h.lower) / h.span * bins));
            }
#pragma unroll
            for (int p = 0; p < kPix; ++p)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (in_lds) atomicAdd(&lh[c * h.bins + bin[p][c]], inc[p][c]);
                else if (inc[p][c]) atomicAdd(h.counts + (size_t)c * h.bins + bin[p][c], 1ull);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int off = 32; off > 0; off >>= 1) {
            under[c] += __shfl_xor(under[c], off, 64);
            over[c] += __shfl_xor(over[c], off, 64);
            nonfinite[c] += __shfl_xor(nonfinite[c], off, 64);
            kmin[c] = min(kmin[c], (uint32_t)__shfl_xor((int)kmin[c], off, 64));
            kmax[c] = max(kmax[c], (uint32_t)__shfl_xor((int)kmax[c], off, 64));
        }
    }
    if (in_lds && h.partial) {
        // The workgroup's results into its own row of the partial array
        // (plain stores): bins, then per channel under / over / nonfinite
        // (64-bit) and the order keys of min / max.
        __shared__ uint32_t wres[kHistThreads / 64][15];
        if ((threadIdx.x & 63) == 0) {
            uint32_t* r = wres[threadIdx.x >> 6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r[3 * c] = under[c]; r[3 * c + 1] = over[c]; r[3 * c + 2] = nonfinite[c];
                r[9 + 2 * c] = kmin[c]; r[10 + 2 * c] = kmax[c];
            }
        }
        __syncthreads();
        uint32_t* row = h.partial + (size_t)blockIdx.x * h.pstride;
        for (uint32_t i = threadIdx.x; i < nb; i += kHistThreads) row[i] = lh[i];
        if (threadIdx.x < 15) {
            const uint32_t q = threadIdx.x;
            uint64_t t = 0;
            uint32_t lo = ~0u, hi = 0u;
            for (int w = 0; w < kHistThreads / 64; ++w) {
                t += wres[w][q];
                lo = min(lo, wres[w][q]);
                hi = max(hi, wres[w][q]);
            }
            uint32_t* tail = row + nb;   // 9 tallies (two words each), then 6 keys
            if (q < 9) { tail[2 * q] = (uint32_t)t; tail[2 * q + 1] = (uint32_t)(t >> 32); }
            else tail[18 + (q - 9)] = ((q - 9) & 1) ? hi : lo;
        }
        return;
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (under[c]) atomicAdd(h.tallies + 3 * c, (unsigned long long)under[c]);
            if (over[c]) atomicAdd(h.tallies + 3 * c + 1, (unsigned long long)over[c]);
            if (nonfinite[c]) atomicAdd(h.tallies + 3 * c + 2, (unsigned long long)nonfinite[c]);
            if (kmin[c] != ~0u) atomicMin(h.extrema + 2 * c, kmin[c]);
            if (kmax[c] != 0u) atomicMax(h.extrema + 2 * c + 1, kmax[c]);
        }
    }
    if (in_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += kHistThreads)
            if (lh[i]) atomicAdd(h.counts + i, (unsigned long long)lh[i]);
    }
}


// ---------------------------------------------------------- local statistics --
//
// The reference's operation order, so FP64 results are equal bit for bit:
// per (row, column) a Welford pass over the row's window in increasing x
// (count, mean += delta / count, m2 += delta * (v - mean), non-finite
// samples skipped), then per output pixel Chan's pairwise merge of those row
// moments in increasing y.  A workgroup owns a 32 x 16 output tile: its
// threads first form the row moments of the tile's rows grown by r into LDS
// (each read by up to 2r + 1 outputs), then merge them.

constexpr int kLsThreads = 256;
constexpr int kLsTw = 32, kLsTh = 16, kLsMaxR = 8;
constexpr int kLsRows = kLsTh + 2 * kLsMaxR;

struct Moment {
    uint32_t n;
    double mean, m2;
};

__device__ __forceinline__ void moment_add(Moment& m, float v) {
    if (!__builtin_isfinite(v)) return;
    ++m.n;
    const double d = (double)v - m.mean;
    m.mean += d / (double)m.n;
    m.m2 += d * ((double)v - m.mean);
}

__device__ __forceinline__ void moment_merge(Moment& a, const Moment& b) {
    if (!b.n) return;
    if (!a.n) { a = b; return; }
    const uint32_t total = a.n + b.n;
    const double d = b.mean - a.mean;
    a.m2 += b.m2 + d * d * ((double)a.n * (double)b.n / (double)total);
    a.mean += d * ((double)b.n / (double)total);
    a.n = total;
}

__global__ void __launch_bounds__(kLsThreads) local_stats_kernel(const LocalStatsLaunch L) {
    __shared__ uint32_t sn[kLsRows * kLsTw * 3];
    __shared__ double smean[kLsRows * kLsTw * 3];
    __shared__ double sm2[kLsRows * kLsTw * 3];
    const int r = (int)L.radius;
    const int ox0 = (int)L.output.x + (int)blockIdx.x * kLsTw;   // tile origin, level coordinates
    const int oy0 = (int)L.output.y + (int)blockIdx.y * kLsTh;
    const int ox_end = min(ox0 + kLsTw, (int)(L.output.x + L.output.width));
    const int oy_end = min(oy0 + kLsTh, (int)(L.output.y + L.output.height));
    const int ix0 = (int)L.image.x, ix1 = (int)(L.image.x + L.image.width) - 1;
    const int iy0 = (int)L.image.y, iy1 = (int)(L.image.y + L.image.height) - 1;
    const int row0 = max(iy0, oy0 - r), row1 = min(iy1, oy_end - 1 + r);   // rows with moments
    const int rows = row1 - row0 + 1;
    // Row moments: (row, column) items, channels together.
    for (int i = threadIdx.x; i < rows * kLsTw; i += kLsThreads) {
        const int ry = i / kLsTw, cx = i % kLsTw;
        const int x = ox0 + cx, y = row0 + ry;
        if (x >= ox_end) continue;
        const int left = max(ix0, x - r), right = min(ix1, x + r);
        const float* src = L.in + ((size_t)(y - (int)L.needed.y) * L.in_stride + (size_t)(left - (int)L.needed.x)) * 3;
        Moment m[3] = {{0, 0.0, 0.0}, {0, 0.0, 0.0}, {0, 0.0, 0.0}};
        for (int xx = left; xx <= right; ++xx, src += 3)
            for (int c = 0; c < 3; ++c) moment_add(m[c], src[c]);
        for (int c = 0; c < 3; ++c) {
            const int k = (ry * kLsTw + cx) * 3 + c;
            sn[k] = m[c].n;
            smean[k] = m[c].mean;
            sm2[k] = m[c].m2;
        }
    }
    __syncthreads();
    const double nan = __builtin_nan("");
    for (int i = threadIdx.x; i < kLsTh * kLsTw; i += kLsThreads) {
        const int ty = i / kLsTw, cx = i % kLsTw;
        const int x = ox0 + cx, y = oy0 + ty;
        if (x >= ox_end || y >= oy_end) continue;
        const int top = max(iy0, y - r), bottom = min(iy1, y + r);
        const size_t o = ((size_t)(y - (int)L.output.y) * L.output.width + (size_t)(x - (int)L.output.x)) * 3;
        for (int c = 0; c < 3; ++c) {
            Moment t = {0, 0.0, 0.0};
            for (int yy = top; yy <= bottom; ++yy) {
                const int k = ((yy - row0) * kLsTw + cx) * 3 + c;
                const Moment b = {sn[k], smean[k], sm2[k]};
                moment_merge(t, b);
            }
            L.count[o + c] = t.n;
            if (t.n) {
                const double v = t.m2 / (double)t.n;
                L.mean[o + c] = t.mean;
                L.variance[o + c] = 0.0 < v ? v : 0.0;   // std::max(0.0, v)
            } else {
                L.mean[o + c] = nan;
                L.variance[o + c] = nan;
            }
        }
    }
}

}  // namespace

// Sum of the workgroups' partial rows into the caller's counters (which
// accumulate across calls): thread (o, chunk) sums output o (a bin, a tally,
// or the min / max of an extremum key) over rows [16 chunk, 16 chunk + 16)
// with eight loads in flight, then adds it with one atomic.
constexpr uint32_t kMergeRows = 16;
__global__ void __launch_bounds__(256) histogram_merge(const HistogramLaunch h, uint32_t rows) {
    const uint32_t nb = 3 * h.bins;
    const uint32_t o = blockIdx.x * 256 + threadIdx.x;
    const uint32_t g0 = blockIdx.y * kMergeRows, g1 = min(rows, g0 + kMergeRows);
    if (o >= nb + 15 || g0 >= g1) return;
    const uint32_t* base = h.partial + (size_t)g0 * h.pstride;
    auto word = [&](uint32_t g, uint32_t w) { return base[(size_t)g * h.pstride + w]; };
    if (o < nb) {
        uint64_t t = 0;
        uint32_t g = 0;
        for (; g + 8 <= g1 - g0; g += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = word(g + u, o);
#pragma unroll
            for (int u = 0; u < 8; ++u) t += v[u];
        }
        for (; g < g1 - g0; ++g) t += word(g, o);
        if (t) atomicAdd(h.counts + o, (unsigned long long)t);
    } else if (o < nb + 9) {
        const uint32_t q = o - nb;
        uint64_t t = 0;
        for (uint32_t g = 0; g < g1 - g0; ++g)
            t += (uint64_t)word(g, nb + 2 * q) | ((uint64_t)word(g, nb + 2 * q + 1) << 32);
        if (t) atomicAdd(h.tallies + q, (unsigned long long)t);
    } else {
        const uint32_t q = o - nb - 9;   // channel q / 2, min (even) or max (odd)
        uint32_t lo = ~0u, hi = 0u;
        for (uint32_t g = 0; g < g1 - g0; ++g) {
            const uint32_t k = word(g, nb + 18 + q);
            lo = min(lo, k);
            hi = max(hi, k);
        }
        if (q & 1) { if (hi != 0u) atomicMax(h.extrema + q, hi); }
        else if (lo != ~0u) atomicMin(h.extrema + q, lo);
    }
}

hipError_t launch_histogram(const HistogramLaunch& h0, hipStream_t stream) {
    HistogramLaunch h = h0;
    const size_t n = (size_t)h.width * h.height;
    if (!n) return hipSuccess;
    if (h.bins <= kLdsBins && h.height >= 64) {
        // Workgroups write partial rows that a second kernel sums: no
        // contention on the caller's counters, so the grid can hold four
        // waves per SIMD (1024 workgroups of four waves: one per SIMD slot).
        // 24 MP, 256 bins: 130 µs + 8.5 µs merge, against 0.245 ms with 512
        // workgroups adding into the counters.  Rows of 3 bins + 24 words;
        // the partial array stays under about a tenth of the image bytes.
        h.pstride = (3 * h.bins + 24 + 3) & ~3u;
        const size_t cap = std::max<size_t>(64, (size_t)n * 12 / 10 / ((size_t)h.pstride * 4));
        const unsigned blocks = (unsigned)std::min<size_t>({1024, cap, (size_t)h.height});
        void* p = nullptr;
        hipError_t e = hipMallocAsync(&p, (size_t)blocks * h.pstride * 4, stream);
        if (e != hipSuccess) return e;
        h.partial = static_cast<uint32_t*>(p);
        hipLaunchKernelGGL(histogram_rgb<true>, dim3(blocks), dim3(kHistThreads), 0, stream, h);
        e = hipGetLastError();
        if (e == hipSuccess) {
            const dim3 grid((3 * h.bins + 15 + 255) / 256, (blocks + kMergeRows - 1) / kMergeRows);
            hipLaunchKernelGGL(histogram_merge, grid, dim3(256), 0, stream, h, blocks);
            e = hipGetLastError();
        }
        const hipError_t f = hipFreeAsync(p, stream);
        return e != hipSuccess ? e : f;
    }
    h.partial = nullptr;
    // At least eight pixels per thread, and at most 512 workgroups of 256
    // threads: each adds its bins to the same global counters, and the waves
    // of a workgroup share its LDS bins (on a 24 MP frame 1024 workgroups ran
    // 0.34 ms against 0.255, 256 workgroups of 512 threads 0.28 ms, of 1024
    // threads 0.35 ms; per-wave bin copies gained nothing).
    const size_t want = (n + (size_t)kHistThreads * 8 - 1) / ((size_t)kHistThreads * 8);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>({want, 512, (size_t)h.height}));
    if (h.bins <= kLdsBins) hipLaunchKernelGGL(histogram_rgb<true>, dim3(blocks), dim3(kHistThreads), 0, stream, h);
    else hipLaunchKernelGGL(histogram_rgb<false>, dim3(blocks), dim3(kHistThreads), 0, stream, h);
    return hipGetLastError();
}

hipError_t launch_local_stats(const LocalStatsLaunch& L, hipStream_t stream) {
    if (!L.output.width || !L.output.height) return hipSuccess;
    const dim3 grid((L.output.width + kLsTw - 1) / kLsTw, (L.output.height + kLsTh - 1) / kLsTh);
    hipLaunchKernelGGL(local_stats_kernel, grid, dim3(kLsThreads), 0, stream, L);
    return hipGetLastError();
}

}  // namespace rawops
