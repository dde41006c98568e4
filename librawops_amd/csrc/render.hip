// Fused Bayer -> RGB render kernels for gfx950.
//
// One workgroup renders one native tile of TW x TH sensor pixels entirely in
// LDS: normalised raw samples (+halo) are staged once, the demosaic runs stage
// by stage over the tile and its shrinking halos, the native op program (white
// balance, exposure, camera matrix, ...) runs in registers, then the tile is
// either written (mip 0) or reduced over 2x2 / 4x4 footprints in FP64 and
// finished with the reduced op program (mip 1/2).
//
// Menon (DDFAPD) arithmetic is FP64 in the reference's exact operation order
// (MenonDemosaic.cpp:89-174); this file is compiled with -ffp-contract=off so
// the result is bit-identical to the reference.  Every stage after the raw
// load works on a checkerboard: the directional green, the gradients and the
// classifier exist only at non-green sites, the colour differences only at
// green sites, so those planes are stored compacted (one value per 2 columns).
//
// Tiles that keep their whole 6-pixel support inside the active area take the
// INTERIOR path (no mirroring, no zero extension); tiles near the edge take the
// BORDER path, which reflects reads at the active-area edge and clips
// computation to the area exactly as the reference's clipped planes do.
#include <atomic>

#include "render_tile.hpp"
#include "render_launch.hpp"
#include "stream_launch.hpp"

namespace rawops {

namespace {

template <int DEMOSAIC, int TW, int TH>
__global__ void __launch_bounds__(kRenderThreads) render_kernel(const RenderArgs a) {
    __shared__ __attribute__((aligned(16))) char lds[Lds<DEMOSAIC, TW, TH>::BYTES];
    int tx, ty, lx0, ly0;
    tile_origin<TW, TH>(a, (int)blockIdx.x, tx, ty, lx0, ly0);
    const SensorArgs s = a.s;
    render_any_tile<DEMOSAIC, TW, TH, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
}

uint32_t site_of_host(const SensorArgs& s, int x, int y) {
    return ((((uint32_t)y & 1u) + s.phase_y) & 1u) * 2u + ((((uint32_t)x & 1u) + s.phase_x) & 1u);
}

} // namespace

// Menon mip-0 render path (rawops_set_menon_path): 0 automatic, 1 streaming
// wherever the area allows it, 2 tile kernel only, 3 as 1 with the Menon
// stages of linear chains in FP32 (approximate); rows per streaming
// wavefront (0: 16).
std::atomic<uint32_t> g_menon_path{0}, g_stream_band{0};

namespace {

// The streaming kernel wants enough wavefronts to fill the chip: viewports
// smaller than about 1536 strips of 112 x 16 (2.75 MP) render faster as
// tiles, whose 512-thread workgroups split them finer.
int stream_band_rows();

bool use_stream(const RenderArgs& a) {
    const SensorArgs& S = a.s;
    if (S.aw < 16 || S.ah < 16 || a.ostride > (1u << 23)) return false;
    const uint32_t path = g_menon_path.load(std::memory_order_relaxed);
    if (path == 2) return false;
    if (path == 1 || path == 3) return true;
    const int band = stream_band_rows() ? stream_band_rows() : 16;
    return (long)((a.vw + 111) / 112) * ((a.vh + band - 1) / band) >= 1536;
}

int stream_band_rows() { return (int)g_stream_band.load(std::memory_order_relaxed); }

// The FP32 output's op program (native ops, then the reduced ops before the
// tap) as a linear chain [gains] [multiplier] [3x3] in this order, each at
// most once: its parameters go into the streaming kernel's arguments.
// Returns the kernel's op form (STREAM_OPS_LISTS when it is not one).
int linear_chain(const RenderLaunch& L, uint32_t tap, StreamArgs& t) {
    if (!L.native_host || !L.reduced_host || L.out8) return STREAM_OPS_LISTS;
    DevOp seq[2 * kMaxOps];
    uint32_t n = 0;
    for (uint32_t i = 0; i < L.native_host->n; ++i) seq[n++] = L.native_host->op[i];
    const uint32_t nr = std::min(tap, L.reduced_host->n);
    for (uint32_t i = 0; i < nr; ++i) seq[n++] = L.reduced_host->op[i];
    t.gains[0] = t.gains[1] = t.gains[2] = 1.0f;
    t.mult = 1.0f;
    int form = STREAM_OPS_LINEAR;
    uint32_t next = 0;   // the next kind allowed: 0 gains, 1 multiplier, 2 matrix
    for (uint32_t i = 0; i < n; ++i) {
        const DevOp& o = seq[i];
        if (o.kind == OP_GAINS && next <= 0) {
            t.gains[0] = o.v[0]; t.gains[1] = o.v[1]; t.gains[2] = o.v[2];
            form |= STREAM_LIN_GAINS; next = 1;
        } else if (o.kind == OP_SCALE && next <= 1) {
            t.mult = o.v[0];
            form |= STREAM_LIN_GAINS; next = 2;
        } else if (o.kind == OP_MATRIX && next <= 2) {
            for (int k = 0; k < 9; ++k) t.m[k] = o.v[k];
            form |= STREAM_LIN_MATRIX; next = 3;
        } else {
            return STREAM_OPS_LISTS;
        }
    }
    return form;
}

template <int DEMOSAIC, int TW, int TH>
hipError_t launch(const RenderLaunch& L, const OpList* native, const OpList* reduced, hipStream_t stream) {
    RenderArgs a;
    a.s = L.sensor;
    a.native = native;
    a.reduced = reduced;
    a.mip = (int)L.mip;
    a.vx = (int)L.viewport.x; a.vy = (int)L.viewport.y;
    a.vw = (int)L.viewport.width; a.vh = (int)L.viewport.height;
    const int scale = 1 << a.mip;
    if (a.mip == 0) { a.ox0 = a.vx; a.oy0 = a.vy; }
    else { a.ox0 = a.s.ax + a.vx * scale; a.oy0 = a.s.ay + a.vy * scale; }
    const int lw = TW / scale, lh = TH / scale;
    a.tiles_x = (a.vw + lw - 1) / lw;
    const int tiles_y = (a.vh + lh - 1) / lh;
    a.out = L.out;
    a.out8 = L.out8;
    a.ostride = L.ostride;
    a.flags = L.flags;
    a.tap = 0xffffffffu;
    a.decisions = L.decisions;
    a.dtiles = decision_tiles((uint32_t)a.vw);
    const unsigned blocks = (unsigned)a.tiles_x * (unsigned)tiles_y;
    if (!blocks) return hipSuccess;
    const SensorArgs& S = a.s;
    StreamArgs t{};
    const int ops = DEMOSAIC == RAWOPS_DEMOSAIC_MENON && a.mip == 0 ? linear_chain(L, a.tap, t) : STREAM_OPS_LISTS;
    // Recording forwards stream only with linear chains (the recording
    // forms are compiled for those).
    if (DEMOSAIC == RAWOPS_DEMOSAIC_MENON && a.mip == 0 && (!a.decisions || ops != STREAM_OPS_LISTS) &&
        use_stream(a)) {
        // The streaming kernel renders the whole viewport (its edge form
        // handles the strips that reach the edge of the active area).
        t.samples = S.samples;
        t.stride = S.stride;
        t.ax = S.ax; t.ay = S.ay; t.aw = S.aw; t.ah = S.ah;
        t.X0 = a.vx; t.X1 = a.vx + a.vw;
        t.Y0 = a.vy; t.Y1 = a.vy + a.vh;
        t.vx = a.vx; t.vy = a.vy;
        t.band_rows = stream_band_rows();
        for (int q = 0; q < 2; ++q)
            for (int c = 0; c < 2; ++c) {
                const uint32_t site = site_of_host(S, t.X0 + c, t.Y0 + q);
                t.blk[q][c] = S.black[site];
                t.rcp[q][c] = S.rcp[site];
                t.den[q][c] = S.denom[site];
            }
        t.native = native;
        t.reduced = reduced;
        t.tap = a.tap;
        t.out = a.out;
        t.out8 = a.out8;
        t.ostride = a.ostride;
        t.flags = a.flags;
        t.decisions = a.decisions;
        t.dtiles = a.dtiles;
        t.fp32 = g_menon_path.load(std::memory_order_relaxed) == 3 ? 1 : 0;
        const int n0 = (int)(((S.ng_parity >> ((uint32_t)t.Y0 & 1u)) & 1u) ^ ((uint32_t)t.X0 & 1u));
        const uint32_t site_ng = site_of_host(S, t.X0 + n0, t.Y0);
        const int red0 = ((S.colors >> (2u * site_ng)) & 3u) == 0 ? 1 : 0;
        return launch_menon_stream(t, n0, red0, ops, stream);
    }
    hipLaunchKernelGGL((render_kernel<DEMOSAIC, TW, TH>), dim3(blocks), dim3(kRenderThreads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_render(const RenderLaunch& L, const OpList* native, const OpList* reduced, hipStream_t stream) {
    if (L.demosaic == RAWOPS_DEMOSAIC_MENON) return launch<RAWOPS_DEMOSAIC_MENON, 64, 32>(L, native, reduced, stream);
    return launch<RAWOPS_DEMOSAIC_BILINEAR, 64, 32>(L, native, reduced, stream);
}

}  // namespace rawops
