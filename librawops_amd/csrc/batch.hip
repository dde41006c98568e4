// Batched heterogeneous render: many requests (different sources, viewports,
// mip levels, demosaicers and op programs) in one persistent launch.
//
// The batch is flattened into a work list of (request, tile) items; each
// request owns a contiguous range of item numbers.  A grid sized to fill the
// chip once claims items from a device counter until the list is exhausted,
// runs the same fused tile as the one-request kernel, and the workgroup that
// finishes a request's last tile stamps its completion time (s_memrealtime,
// 100 MHz).  Counter reset + persistent kernel are captured in a hipGraph and
// replayed per batch.  The host-side autotuner times the compiled launch
// configurations (tile shape x resident workgroups per CU) on the actual
// batch and keeps the fastest.
#include "render_tile_legacy.hpp"
#include "batch_launch.hpp"

namespace rawops {
namespace {

constexpr int cmax(int a, int b) { return a > b ? a : b; }
// One LDS allocation large enough for both demosaicers at the 64x32 tile.
constexpr int kBatchLds = cmax(LegacyLds<RAWOPS_DEMOSAIC_MENON, 64, 32>::BYTES, LegacyLds<RAWOPS_DEMOSAIC_BILINEAR, 64, 32>::BYTES);

struct BatchState {
    uint32_t next;
    uint32_t pad;
    unsigned long long t0;
};

__global__ void batch_reset(BatchState* st, const BatchItem* items, int n, uint32_t* remaining) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->next = 0;
        st->t0 = wall_clock64();
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) remaining[i] = items[i].tiles;
}

template <int TW, int TH>
__device__ __forceinline__ void run_tile(const RenderArgs& a, int demosaic, int tile, char* lds) {
    int tx, ty, lx0, ly0;
    tile_origin<TW, TH>(a, tile, tx, ty, lx0, ly0);
    const SensorArgs s = a.s;
    if (demosaic == RAWOPS_DEMOSAIC_MENON) {
        const bool interior = tx - 6 >= s.ax && ty - 6 >= s.ay && tx + TW + 6 <= s.ax + s.aw && ty + TH + 6 <= s.ay + s.ah;
        if (interior) render_tile_legacy<RAWOPS_DEMOSAIC_MENON, TW, TH, false>(a, s, lds, tx, ty, lx0, ly0);
        else render_tile_legacy<RAWOPS_DEMOSAIC_MENON, TW, TH, true>(a, s, lds, tx, ty, lx0, ly0);
    } else {
        const bool interior = tx - 1 >= s.ax && ty - 1 >= s.ay && tx + TW + 1 <= s.ax + s.aw && ty + TH + 1 <= s.ay + s.ah;
        if (interior) render_tile_legacy<RAWOPS_DEMOSAIC_BILINEAR, TW, TH, false>(a, s, lds, tx, ty, lx0, ly0);
        else render_tile_legacy<RAWOPS_DEMOSAIC_BILINEAR, TW, TH, true>(a, s, lds, tx, ty, lx0, ly0);
    }
}

template <int TW, int TH>
__global__ void __launch_bounds__(kThreads) batch_kernel(const BatchItem* __restrict__ items, int n, int total,
                                                         BatchState* st, uint32_t* remaining,
                                                         unsigned long long* done) {
    __shared__ __attribute__((aligned(16))) char lds[kBatchLds];
    __shared__ int claim;
    for (;;) {
        if (threadIdx.x == 0) claim = (int)atomicAdd(&st->next, 1u);
        __syncthreads();
        const int item = claim;
        __syncthreads();
        if (item >= total) break;
        int lo = 0, hi = n - 1;                     // request owning this item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].first <= item) lo = mid; else hi = mid - 1;
        }
        const BatchItem& q = items[lo];
        RenderArgs a;
        a.decisions = nullptr;
        a.dtiles = 0;
        a.s = q.sensor;
        a.native = q.native;
        a.reduced = q.reduced;
        a.mip = q.mip;
        a.vx = q.vx; a.vy = q.vy; a.vw = q.vw; a.vh = q.vh;
        a.ox0 = q.ox0; a.oy0 = q.oy0;
        a.tiles_x = q.tiles_x;
        a.out = q.out;
        a.out8 = q.out8;
        a.ostride = q.ostride;
        a.flags = q.flags;
        a.tap = q.tap;
        run_tile<TW, TH>(a, q.demosaic, item - q.first, lds);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicSub(&remaining[lo], 1u) == 1u) done[lo] = wall_clock64();
        }
    }
}

// One workgroup per (request, tile) item, the render kernel's 512-thread
// tile: no claim loop.  The workgroup that finishes a request's last tile
// stamps its completion time, as in the persistent kernel.
__global__ void __launch_bounds__(kRenderThreads) __attribute__((amdgpu_waves_per_eu(6)))
batch_tiles(const BatchItem* __restrict__ items, int n, uint32_t* remaining, unsigned long long* done) {
    __shared__ __attribute__((aligned(16))) char lds[cmax(Lds<RAWOPS_DEMOSAIC_MENON, 64, 32>::BYTES,
                                                          Lds<RAWOPS_DEMOSAIC_BILINEAR, 64, 32>::BYTES)];
    const int item = (int)blockIdx.x;
    int lo = 0, hi = n - 1;                         // request owning this item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].first <= item) lo = mid; else hi = mid - 1;
    }
    const BatchItem& q = items[lo];
    RenderArgs a;
    a.decisions = nullptr;
    a.dtiles = 0;
    a.s = q.sensor;
    a.native = q.native;
    a.reduced = q.reduced;
    a.mip = q.mip;
    a.vx = q.vx; a.vy = q.vy; a.vw = q.vw; a.vh = q.vh;
    a.ox0 = q.ox0; a.oy0 = q.oy0;
    a.tiles_x = q.tiles_x;
    a.out = q.out;
    a.out8 = q.out8;
    a.ostride = q.ostride;
    a.flags = q.flags;
    a.tap = q.tap;
    int tx, ty, lx0, ly0;
    tile_origin<64, 32>(a, item - q.first, tx, ty, lx0, ly0);
    const SensorArgs s = a.s;
    if (q.demosaic == RAWOPS_DEMOSAIC_MENON)
        render_any_tile<RAWOPS_DEMOSAIC_MENON, 64, 32, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
    else
        render_any_tile<RAWOPS_DEMOSAIC_BILINEAR, 64, 32, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicSub(&remaining[lo], 1u) == 1u) done[lo] = wall_clock64();
    }
}

template <int TW, int TH>
hipError_t launch_variant(const BatchItem* items, int n, int total, void* state, uint32_t* remaining,
                          unsigned long long* done, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(batch_reset, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1), dim3(256), 0, stream,
                       static_cast<BatchState*>(state), items, n, remaining);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((batch_kernel<TW, TH>), dim3(grid), dim3(kThreads), 0, stream, items, n, total,
                       static_cast<BatchState*>(state), remaining, done);
    return hipGetLastError();
}

template <int TW, int TH>
int occupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_kernel<TW, TH>, kThreads, 0) != hipSuccess) return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace

// Launch configurations the autotuner chooses between: tile shape and the
// number of resident workgroups per CU the persistent grid is sized for
// (0 = as many as fit), or one workgroup per item (-1).  Only the 64x32 tile
// is compiled in.
const BatchShape kBatchShapes[kBatchShapeCount] = {{64, 32, 0}, {64, 32, 1}, {64, 32, -1}};

size_t batch_state_bytes() { return sizeof(BatchState); }

int batch_grid(int shape) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (kBatchShapes[shape].per_cu < 0) return 0;   // grid = item count, set at launch
    const int fit = occupancy<64, 32>();
    const int per = kBatchShapes[shape].per_cu > 0 && kBatchShapes[shape].per_cu < fit ? kBatchShapes[shape].per_cu : fit;
    return cus * per;
}

hipError_t launch_batch(int shape, const BatchItem* items, int n, int total, void* state, uint32_t* remaining,
                        unsigned long long* done, int grid, hipStream_t stream) {
    if (kBatchShapes[shape].per_cu < 0) {
        hipLaunchKernelGGL(batch_reset, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1), dim3(256), 0, stream,
                           static_cast<BatchState*>(state), items, n, remaining);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || total <= 0) return e;
        hipLaunchKernelGGL(batch_tiles, dim3(total), dim3(kRenderThreads), 0, stream, items, n, remaining, done);
        return hipGetLastError();
    }
    return launch_variant<64, 32>(items, n, total, state, remaining, done, grid, stream);
}

}  // namespace rawops
