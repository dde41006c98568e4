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
#include <cstdlib>
#include <mutex>

#include "render_tile.hpp"
#include "render_launch.hpp"
#include "stream_launch.hpp"

namespace rawops {

namespace {

template <int DEMOSAIC, int TW, int TH>
__global__ void __launch_bounds__(kRenderThreads) render_kernel(const RenderArgs a) {
    __shared__ __attribute__((aligned(16))) char lds[Lds<DEMOSAIC, TW, TH>::BYTES];
    int tx, ty, lx0, ly0;
    tile_origin<TW, TH>(a, border_tile(a, (int)blockIdx.x), tx, ty, lx0, ly0);
    const SensorArgs s = a.s;
    render_any_tile<DEMOSAIC, TW, TH, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
}

int ceil_div(int a, int b) { return a >= 0 ? (a + b - 1) / b : -((-a) / b); }
int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

uint32_t site_of_host(const SensorArgs& s, int x, int y) {
    return ((((uint32_t)y & 1u) + s.phase_y) & 1u) * 2u + ((((uint32_t)x & 1u) + s.phase_x) & 1u);
}

// The streaming kernel renders the interior of Menon mip-0 viewports unless
// RAWOPS_MENON_STREAM=0 (tile kernel only); RAWOPS_STREAM_BAND sets its rows
// per wavefront (a multiple of four; default 16).
bool stream_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("RAWOPS_MENON_STREAM");
        return !(v && v[0] == '0');
    }();
    return on;
}
int stream_band_rows() {
    static const int rows = [] {
        const char* v = std::getenv("RAWOPS_STREAM_BAND");
        const int r = v ? std::atoi(v) : 0;
        return r >= 4 && r <= 4096 ? r & ~3 : 0;
    }();
    return rows;
}

// The border tiles of a split render run on a library-owned stream beside
// the caller's (default priority, one per device): forked after the work
// already on the caller's stream and joined back before anything after the
// render, so the caller sees one stream-ordered operation.  Fork and join
// reuse two events per device; the mutex keeps concurrent callers' record /
// wait pairs from interleaving.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool ready = false;
};
constexpr int kMaxSideDevices = 64;
std::mutex g_side_mutex;
SideStream g_side[kMaxSideDevices];

SideStream* side_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxSideDevices) return nullptr;
    SideStream& s = g_side[dev];
    if (!s.ready) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.join, hipEventDisableTiming) != hipSuccess)
            return nullptr;
        s.ready = true;
    }
    return &s;
}

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
    a.tiles_y = tiles_y;
    a.bx0 = a.bx1 = a.by0 = a.by1 = 0;
    unsigned blocks = (unsigned)a.tiles_x * (unsigned)tiles_y;
    if (!blocks) return hipSuccess;
    if (DEMOSAIC == RAWOPS_DEMOSAIC_MENON && a.mip == 0 && !a.decisions && stream_enabled()) {
        // Tiles whose 6-pixel support lies inside the area form a rectangle
        // of the tile grid: the streaming kernel renders it, the tile kernel
        // the ring of tiles around it.
        const SensorArgs& S = a.s;
        const int i0 = std::max(0, ceil_div(S.ax + 6 - a.vx, TW));
        const int i1 = std::min(a.tiles_x, floor_div(S.ax + S.aw - 6 - TW - a.vx, TW) + 1);
        const int j0 = std::max(0, ceil_div(S.ay + 6 - a.vy, TH));
        const int j1 = std::min(tiles_y, floor_div(S.ay + S.ah - 6 - TH - a.vy, TH) + 1);
        if (i1 - i0 >= 2 && j1 - j0 >= 2) {
            StreamArgs t{};
            t.samples = S.samples;
            t.stride = S.stride;
            t.ax = S.ax; t.ay = S.ay; t.aw = S.aw; t.ah = S.ah;
            t.X0 = a.vx + i0 * TW;
            t.X1 = std::min(a.vx + i1 * TW, a.vx + a.vw);
            t.Y0 = a.vy + j0 * TH;
            t.Y1 = std::min(a.vy + j1 * TH, a.vy + a.vh);
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
            const int n0 = (int)(((S.ng_parity >> ((uint32_t)t.Y0 & 1u)) & 1u) ^ ((uint32_t)t.X0 & 1u));
            const int xng = t.X0 + n0;
            const uint32_t site_ng = site_of_host(S, xng, t.Y0);
            const int red0 = ((S.colors >> (2u * site_ng)) & 3u) == 0 ? 1 : 0;
            const int ops = linear_chain(L, a.tap, t);
            a.bx0 = i0; a.bx1 = i1; a.by0 = j0; a.by1 = j1;
            blocks -= (unsigned)((i1 - i0) * (j1 - j0));
            if (!blocks) return launch_menon_stream(t, n0, red0, ops, stream);
            std::lock_guard<std::mutex> lock(g_side_mutex);
            SideStream* side = side_stream();
            if (!side) {
                hipError_t e = launch_menon_stream(t, n0, red0, ops, stream);
                if (e != hipSuccess) return e;
            } else {
                hipError_t e = hipEventRecord(side->fork, stream);
                if (e == hipSuccess) e = hipStreamWaitEvent(side->stream, side->fork, 0);
                if (e != hipSuccess) return e;
                hipLaunchKernelGGL((render_kernel<DEMOSAIC, TW, TH>), dim3(blocks), dim3(kRenderThreads), 0,
                                   side->stream, a);
                e = hipGetLastError();
                if (e == hipSuccess) e = hipEventRecord(side->join, side->stream);
                if (e == hipSuccess) e = launch_menon_stream(t, n0, red0, ops, stream);
                const hipError_t j = hipStreamWaitEvent(stream, side->join, 0);
                return e != hipSuccess ? e : j;
            }
        }
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
