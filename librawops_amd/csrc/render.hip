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
#include "render_tile.hpp"
#include "render_launch.hpp"

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
    hipLaunchKernelGGL((render_kernel<DEMOSAIC, TW, TH>), dim3(blocks), dim3(kRenderThreads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_render(const RenderLaunch& L, const OpList* native, const OpList* reduced, hipStream_t stream) {
    if (L.demosaic == RAWOPS_DEMOSAIC_MENON) return launch<RAWOPS_DEMOSAIC_MENON, 64, 32>(L, native, reduced, stream);
    return launch<RAWOPS_DEMOSAIC_BILINEAR, 64, 32>(L, native, reduced, stream);
}

}  // namespace rawops
