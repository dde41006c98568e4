// The fused render tile (demosaic + op program + mip reduction + stores),
// shared by the one-request render kernel and the batched launch.
#pragma once

#include "device_common.hpp"
#include "menon_fwd.hpp"
#include "raw_tile.hpp"

namespace rawops {

constexpr int kRenderThreads = 512;  // render and batch workgroups

struct RenderArgs {
    SensorArgs s;
    const OpList* __restrict__ native;   // device copies owned by the plan
    const OpList* __restrict__ reduced;
    int mip;                 // 0, 1, 2
    int vx, vy, vw, vh;      // viewport in level coordinates
    int ox0, oy0;            // native origin of level pixel (vx, vy)
    int tiles_x;
    float* out;              // may be null
    uint8_t* out8;           // may be null
    uint32_t ostride;        // output row stride in pixels
    uint32_t* flags;         // may be null
    uint32_t tap;            // reduced ops applied before the FP32 output; the rest precede the u8 output
    uint32_t* decisions;     // Menon mip 0: classifier decisions of the viewport, one bit per
                             // non-green site, 32 words per tile (word tile 32 + row, bit x / 2
                             // within the tile, 1 = horizontal); may be null
    uint32_t dtiles;         // tiles per viewport row
};

// Tile-relative x parity of the non-green sites in even tile rows for a
// render whose tiles start at native (ox0 + i TW, oy0 + j TH), TW, TH even.
__host__ __device__ inline int tile_n0(const SensorArgs& s, int ox0, int oy0) {
    return (int)(((s.ng_parity >> ((uint32_t)oy0 & 1u)) & 1u) ^ ((uint32_t)ox0 & 1u));
}

// Finish one level pixel: reduced ops up to the tap, FP32 store, remaining
// ops, uint8 store (mip reduction path).
__device__ __forceinline__ void emit_pixel(const RenderArgs& a, int lx, int ly, float r, float g, float b,
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

// ----------------------------------------------------------- the kernel --

template <int TW, int TH>
struct BilinearLds {
    static constexpr int RAW_BYTES = (TW + 2) * (TH + 2) * 4;
    static constexpr int OFF_NATIVE = (RAW_BYTES + 15) / 16 * 16;
    static constexpr int BYTES = OFF_NATIVE + TW * TH * 12;
};

template <int DEMOSAIC, int TW, int TH>
struct Lds;
template <int TW, int TH> struct Lds<RAWOPS_DEMOSAIC_MENON, TW, TH> { static constexpr int BYTES = MenonFwdLds<TW, TH, true>::BYTES; static constexpr int OFF_NATIVE = MenonFwdLds<TW, TH, true>::OFF_NATIVE; };
template <int TW, int TH> struct Lds<RAWOPS_DEMOSAIC_BILINEAR, TW, TH> { static constexpr int BYTES = BilinearLds<TW, TH>::BYTES; static constexpr int OFF_NATIVE = BilinearLds<TW, TH>::OFF_NATIVE; };

// Level pixels per tile side and native tile origin.
template <int TW, int TH>
__device__ __forceinline__ void tile_origin(const RenderArgs& a, int tile, int& tx, int& ty, int& lx0, int& ly0) {
    const int scale = 1 << a.mip;
    const int bx = tile % a.tiles_x, by = tile / a.tiles_x;
    lx0 = a.vx + bx * (TW / scale);
    ly0 = a.vy + by * (TH / scale);
    tx = a.ox0 + bx * TW;
    ty = a.oy0 + by * TH;
}

// One tile on a workgroup of NT threads.  Each thread owns NQ 2x2 cells of
// the native tile: pixel p = (2 j + b) 2 + e of px is (2 qx + e, 2 qy + b) of
// cell threadIdx.x + j NT.
template <int DEMOSAIC, int TW, int TH, int NT, bool BORDER, int N0>
__device__ __forceinline__ void render_tile(const RenderArgs& a, const SensorArgs& s, char* lds, int tx, int ty, int lx0, int ly0) {
    constexpr int NQ = (TW / 2) * (TH / 2) / NT;
    static_assert(NQ * NT == (TW / 2) * (TH / 2), "tile cells must divide evenly over the workgroup");
    constexpr int NPX = NQ * 4;
    float px[NPX][3];
    uint32_t err[NPX];
#pragma unroll
    for (int p = 0; p < NPX; ++p) err[p] = 0;

    if constexpr (DEMOSAIC == RAWOPS_DEMOSAIC_MENON) {
        using F = MenonFwd<TW, TH, NT, N0, BORDER, true>;
        F t(s, lds, tx, ty);
        t.load();
        __syncthreads();
        typename F::Carry carry;
        t.directional(carry);
        __syncthreads();
        t.gradients(carry);
        __syncthreads();
        t.classify();
        __syncthreads();
        // Record the tile's decisions for the adjoint (64x32 tiles only: a
        // tile row's 32 non-green sites are one word, half a wave's ballot).
        if constexpr (TW == 64 && TH == 32) if (a.decisions && a.mip == 0) {
            static_assert(((TW / 2) * TH) % NT == 0, "one decision word per tile row");
#pragma unroll
            for (int j = 0; j < (TW / 2) * TH / NT; ++j) {
                const int i = (int)threadIdx.x + j * NT;
                const int ry = i / (TW / 2), k = i - ry * (TW / 2);
                const int sx = tx + 2 * k + F::nx(ry & 1), sy = ty + ry;
                const bool h = t.hz[i] != 0 && sx < a.vx + a.vw && in_area(s, sx, sy);
                const uint64_t m = __ballot(h);
                const int lane = (int)threadIdx.x & 63;
                if ((lane & 31) == 0 && sy < a.vy + a.vh)
                    a.decisions[((size_t)((ty - a.vy) >> 5) * a.dtiles + (size_t)((tx - a.vx) >> 6)) * 32 + ry] =
                        lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            }
        }
        t.chroma();
        __syncthreads();
        t.finish(px);
    } else {
        float* raw = reinterpret_cast<float*>(lds);
        load_raw<TW, TH, 1, BORDER, NT>(s, tx, ty, raw);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    float* v = px[(2 * j + b) * 2 + e];
                    v[0] = v[1] = v[2] = 0.f;
                    const int rx = 2 * qx + e, ry = 2 * qy + b;
                    if (!BORDER || in_area(s, tx + rx, ty + ry)) bilinear_pixel<TW, TH, BORDER>(s, raw, tx, ty, rx, ry, v[0], v[1], v[2]);
                }
        }
    }

    apply_ops_px(*a.native, 0, a.native->n, px, err);

    if (a.mip == 0) {
        const OpList& R = *a.reduced;
        const uint32_t tap = a.tap < R.n ? a.tap : R.n;
        apply_ops_px(R, 0, tap, px, err);
        if (a.out) {
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const int it = (int)threadIdx.x + j * NT;
                const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int lx = tx + 2 * qx - a.ox0, ly = ty + 2 * qy + b - a.oy0;
                    if (ly >= a.vh || lx >= a.vw) continue;
                    float* o = a.out + ((size_t)ly * a.ostride + (size_t)lx) * 3;
                    const float* p0 = px[(2 * j + b) * 2];
                    const float* p1 = px[(2 * j + b) * 2 + 1];
                    // Streaming stores (non-temporal): the frame's output should not
                    // push its input samples out of the caches.
                    if (lx + 1 < a.vw) {
                        __builtin_nontemporal_store(p0[0], o); __builtin_nontemporal_store(p0[1], o + 1);
                        __builtin_nontemporal_store(p0[2], o + 2); __builtin_nontemporal_store(p1[0], o + 3);
                        __builtin_nontemporal_store(p1[1], o + 4); __builtin_nontemporal_store(p1[2], o + 5);
                    } else {
                        __builtin_nontemporal_store(p0[0], o); __builtin_nontemporal_store(p0[1], o + 1);
                        __builtin_nontemporal_store(p0[2], o + 2);
                    }
                }
            }
        }
        if (a.out8) {
            apply_ops_px(R, tap, R.n, px, err);
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const int it = (int)threadIdx.x + j * NT;
                const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int lx = tx + 2 * qx + e - a.ox0, ly = ty + 2 * qy + b - a.oy0;
                        if (ly >= a.vh || lx >= a.vw) continue;
                        uint8_t* o = a.out8 + ((size_t)ly * a.ostride + (size_t)lx) * 3;
                        const float* v = px[(2 * j + b) * 2 + e];
                        o[0] = quantize_u8(v[0]); o[1] = quantize_u8(v[1]); o[2] = quantize_u8(v[2]);
                    }
            }
        }
    }
    // Errors count only for pixels the request covers: inside the area and
    // inside the viewport (through the mip footprint).
    uint32_t e_all = 0;
#pragma unroll
    for (int p = 0; p < NPX; ++p) e_all |= err[p];
    if (e_all) {
        e_all = 0;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int sx = tx + 2 * qx + e, sy = ty + 2 * qy + b;
                    if (in_area(s, sx, sy) && ((sx - a.ox0) >> a.mip) < a.vw && ((sy - a.oy0) >> a.mip) < a.vh)
                        e_all |= err[(2 * j + b) * 2 + e];
                }
        }
    }

    if (a.mip != 0) {
        // Footprint reduction anchored at the active-area origin
        // (CameraToWorkingNode::render_level RawEngine.cpp:580-601): FP64
        // sums in row-major footprint order, clipped to the area.
        float* native = reinterpret_cast<float*>(lds + Lds<DEMOSAIC, TW, TH>::OFF_NATIVE);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        native[((2 * qy + b) * TW + 2 * qx + e) * 3 + c] = px[(2 * j + b) * 2 + e][c];
        }
        __syncthreads();
        const int scale = 1 << a.mip;
        const int ow = TW / scale, oh = TH / scale;
        const int nx_end = s.ax + s.aw - tx, ny_end = s.ay + s.ah - ty;   // area end, tile-relative
        for (int i = threadIdx.x; i < ow * oh; i += NT) {
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
            emit_pixel(a, lx, ly, (float)(sr / count), (float)(sg / count), (float)(sb / count), e_all);
        }
    }
    if (e_all && a.flags) atomicOr(a.flags, e_all);
}

// Interior or border form of a tile, for either non-green phase.
template <int DEMOSAIC, int TW, int TH, int NT>
__device__ __forceinline__ void render_any_tile(const RenderArgs& a, const SensorArgs& s, char* lds, int tx, int ty,
                                                int lx0, int ly0) {
    const int n0 = tile_n0(s, a.ox0, a.oy0);
    constexpr int H = DEMOSAIC == RAWOPS_DEMOSAIC_MENON ? 6 : 1;
    const bool interior = tx - H >= s.ax && ty - H >= s.ay && tx + TW + H <= s.ax + s.aw && ty + TH + H <= s.ay + s.ah;
    if constexpr (DEMOSAIC == RAWOPS_DEMOSAIC_MENON) {
        if (n0) {
            if (interior) render_tile<DEMOSAIC, TW, TH, NT, false, 1>(a, s, lds, tx, ty, lx0, ly0);
            else render_tile<DEMOSAIC, TW, TH, NT, true, 1>(a, s, lds, tx, ty, lx0, ly0);
        } else {
            if (interior) render_tile<DEMOSAIC, TW, TH, NT, false, 0>(a, s, lds, tx, ty, lx0, ly0);
            else render_tile<DEMOSAIC, TW, TH, NT, true, 0>(a, s, lds, tx, ty, lx0, ly0);
        }
    } else {
        (void)n0;
        if (interior) render_tile<DEMOSAIC, TW, TH, NT, false, 0>(a, s, lds, tx, ty, lx0, ly0);
        else render_tile<DEMOSAIC, TW, TH, NT, true, 0>(a, s, lds, tx, ty, lx0, ly0);
    }
}

}  // namespace rawops
