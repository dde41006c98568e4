// The batch kernel's tile: the fused render tile in its first form (one
// thread per horizontal pixel pair, per-pixel op program, MenonTile stages),
// kept for the persistent batch kernel while the restructured tile
// (render_tile.hpp) is validated there.
#pragma once

#include "render_tile.hpp"

namespace rawops {

template <int DEMOSAIC, int TW, int TH>
struct LegacyLds;
template <int TW, int TH> struct LegacyLds<RAWOPS_DEMOSAIC_MENON, TW, TH> { static constexpr int BYTES = MenonLds<TW, TH>::BYTES; static constexpr int OFF_NATIVE = MenonLds<TW, TH>::OFF_NATIVE; };
template <int TW, int TH> struct LegacyLds<RAWOPS_DEMOSAIC_BILINEAR, TW, TH> { static constexpr int BYTES = BilinearLds<TW, TH>::BYTES; static constexpr int OFF_NATIVE = BilinearLds<TW, TH>::OFF_NATIVE; };

// Finish one level pixel (legacy tile): reduced ops up to the tap, FP32 store, remaining
// ops, uint8 store.
__device__ __forceinline__ void emit_pixel_legacy(const RenderArgs& a, int lx, int ly, float r, float g, float b,
                                           uint32_t& err) {
    const OpList& L = *a.reduced;
    const uint32_t tap = a.tap < L.n ? a.tap : L.n;
    apply_ops(L, 0, tap, r, g, b, err);
    const size_t o = ((size_t)ly * a.ostride + (size_t)lx) * 3;
    if (a.out) { a.out[o] = r; a.out[o + 1] = g; a.out[o + 2] = b; }
    if (a.out8) {
        apply_ops(L, tap, L.n, r, g, b, err);
        a.out8[o] = quantize_u8(r); a.out8[o + 1] = quantize_u8(g); a.out8[o + 2] = quantize_u8(b);
    }
}

template <int DEMOSAIC, int TW, int TH, bool BORDER>
__device__ __forceinline__ void render_tile_legacy(const RenderArgs& a, const SensorArgs& s, char* lds, int tx, int ty, int lx0, int ly0) {
    uint32_t err = 0;
    // Native pixels this thread finishes: pairs of horizontally adjacent
    // pixels, so every lane handles one green and one non-green site.
    constexpr int PAIRS = (TW / 2) * TH;
    static_assert(PAIRS % kThreads == 0, "tile pairs must divide evenly over the workgroup");
    constexpr int PER_THREAD = PAIRS / kThreads;
    float px[PER_THREAD][2][3];

    if constexpr (DEMOSAIC == RAWOPS_DEMOSAIC_MENON) {
        using L = MenonLds<TW, TH>;
        MenonTile<TW, TH, BORDER> t{s, tx, ty, 0,
                                    reinterpret_cast<float*>(lds + L::OFF_RAW),
                                    reinterpret_cast<double*>(lds + L::OFF_GH),
                                    reinterpret_cast<double*>(lds + L::OFF_GV),
                                    reinterpret_cast<double*>(lds + L::OFF_DH),
                                    reinterpret_cast<double*>(lds + L::OFF_DV),
                                    reinterpret_cast<double*>(lds + L::OFF_GREEN),
                                    reinterpret_cast<double*>(lds + L::OFF_GH),
                                    reinterpret_cast<double*>(lds + L::OFF_GV),
                                    reinterpret_cast<uint8_t*>(lds + L::OFF_HZ)};
        // Tile-relative non-green parity for both row parities.
        for (int b = 0; b < 2; ++b) {
            const uint32_t sensor = (s.ng_parity >> (((uint32_t)ty + (uint32_t)b) & 1u)) & 1u;
            t.ngr |= ((sensor ^ ((uint32_t)tx & 1u)) & 1u) << b;
        }
        load_raw<TW, TH, 6, BORDER>(s, tx, ty, t.raw);
        __syncthreads();
        t.directional();
        __syncthreads();
        t.gradients();
        __syncthreads();
        t.classify();
        __syncthreads();
        t.chroma();
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER_THREAD; ++j) {
            const int p = threadIdx.x + j * kThreads;
            const int ry = p / (TW / 2), rx = 2 * (p - ry * (TW / 2));
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                float r = 0.f, g = 0.f, b = 0.f;
                if (t.inside(rx + e, ry)) {
                    t.finish(rx + e, ry, r, g, b);
                    apply_ops(*a.native, r, g, b, err);
                }
                px[j][e][0] = r; px[j][e][1] = g; px[j][e][2] = b;
            }
        }
    } else {
        float* raw = reinterpret_cast<float*>(lds);
        load_raw<TW, TH, 1, BORDER>(s, tx, ty, raw);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER_THREAD; ++j) {
            const int p = threadIdx.x + j * kThreads;
            const int ry = p / (TW / 2), rx = 2 * (p - ry * (TW / 2));
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                float r = 0.f, g = 0.f, b = 0.f;
                if (!BORDER || in_area(s, tx + rx + e, ty + ry)) {
                    bilinear_pixel<TW, TH, BORDER>(s, raw, tx, ty, rx + e, ry, r, g, b);
                    apply_ops(*a.native, r, g, b, err);
                }
                px[j][e][0] = r; px[j][e][1] = g; px[j][e][2] = b;
            }
        }
    }

    if (a.mip == 0) {
#pragma unroll
        for (int j = 0; j < PER_THREAD; ++j) {
            const int p = threadIdx.x + j * kThreads;
            const int ry = p / (TW / 2), rx = 2 * (p - ry * (TW / 2));
            const int ly = ty + ry - a.oy0;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int lx = tx + rx + e - a.ox0;
                if (lx >= a.vw || ly >= a.vh) continue;
                emit_pixel_legacy(a, lx, ly, px[j][e][0], px[j][e][1], px[j][e][2], err);
            }
        }
    } else {
        // Footprint reduction anchored at the active-area origin
        // (CameraToWorkingNode::render_level RawEngine.cpp:580-601): FP64
        // sums in row-major footprint order, clipped to the area.
        float* native = reinterpret_cast<float*>(lds + LegacyLds<DEMOSAIC, TW, TH>::OFF_NATIVE);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER_THREAD; ++j) {
            const int p = threadIdx.x + j * kThreads;
            const int ry = p / (TW / 2), rx = 2 * (p - ry * (TW / 2));
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int c = 0; c < 3; ++c) native[(ry * TW + rx + e) * 3 + c] = px[j][e][c];
        }
        __syncthreads();
        const int scale = 1 << a.mip;
        const int ow = TW / scale, oh = TH / scale;
        const int nx_end = s.ax + s.aw - tx, ny_end = s.ay + s.ah - ty;   // area end, tile-relative
        for (int i = threadIdx.x; i < ow * oh; i += kThreads) {
            const int oy = i / ow, ox = i - oy * ow;
            const int lx = lx0 + ox - a.vx, ly = ly0 + oy - a.vy;
            if (lx >= a.vw || ly >= a.vh) continue;
            double sr = 0, sg = 0, sb = 0;
            int count = 0;
            for (int fy = oy * scale; fy < oy * scale + scale && fy < ny_end; ++fy)
                for (int fx = ox * scale; fx < ox * scale + scale && fx < nx_end; ++fx) {
                    const float* q = native + (fy * TW + fx) * 3;
                    sr += q[0]; sg += q[1]; sb += q[2];
                    ++count;
                }
            emit_pixel_legacy(a, lx, ly, (float)(sr / count), (float)(sg / count), (float)(sb / count), err);
        }
    }
    if (err && a.flags) atomicOr(a.flags, err);
}

}  // namespace rawops
