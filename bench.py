"""Benchmark: 24 MP Menon render on one MI355X.

Workload (the headline configuration): 6048x4024 uint16 Bayer, 14-bit, BGGR
with CFA phase (1, 0), row stride 6080, active area inset 16 px (6016x3992 =
24,015,872 px), per-site black (510, 512, 511, 513), white 16383, seeded
camera->ProPhoto matrix; Menon demosaic -> white balance -> exposure -> 3x3,
FP32 linear output of the full active area.  One step = one render of the
frame (one fused kernel launch) with the source already resident in HBM.

Prints one JSON line (see the field list at the end of main()).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_PER_PX = 14  # 2 B Bayer in + 12 B FP32 RGB out


def cpu_baseline(meta, samples, cam, threads, rows=512):
    """The oracle restatement (C, OpenMP row bands) on a bounded band of the
    same frame: Menon + WB + exposure + 3x3, the fastest reference-legal
    configuration (concurrent band requests)."""
    import numpy as np

    from oracle import oracle as O
    om = O.Bayer(meta["width"], meta["height"], meta["row_stride_samples"], meta["pattern"],
                 meta["cfa_phase_x"], meta["cfa_phase_y"], meta["active_area"], meta["black_levels"],
                 meta["white_levels"])
    ax, ay, aw, ah = meta["active_area"]
    rect = (ax, ay, aw, min(rows, ah))
    m = O.camera_to_working_f32(cam, 0)
    reps, t0 = 0, time.perf_counter()
    while True:
        nat = O.menon_bands(om, samples, rect, threads)
        O.scale3(nat, (1.0, 1.0, 1.0))
        O.scale(nat, O.exp2f(0.0))
        O.mat3(nat, m)
        reps += 1
        el = time.perf_counter() - t0
        if el > 8.0 or reps >= 20:
            break
    px = rect[2] * rect[3] * reps
    return {"value": px / el / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"{reps}x Menon+WB+exposure+3x3 on a {rect[2]}x{rect[3]} band of the same frame, "
                      f"{threads} OpenMP row bands, oracle C restatement (gcc -O2)"}


def compiled_matrix(node):
    """float32 camera->working coefficients the CameraToWorkingNode uses."""
    return node.matrix


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 256 MiB device copy (reference for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 64 * 1024 * 1024  # floats = 256 MiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import librawops_amd as L
    from librawops_amd import synth
    from librawops_amd.graph import DeviceRender, compile_chain

    torch.cuda.set_device(0)
    meta, samples, cam = synth.config2_menon(seed=2)
    md = L.RawMetadata(meta["width"], meta["height"], meta["row_stride_samples"], L.BayerPattern(meta["pattern"]),
                       meta["cfa_phase_x"], meta["cfa_phase_y"], L.Rect(*meta["active_area"]),
                       meta["black_levels"], meta["white_levels"])
    dev_samples = torch.from_numpy(samples.view(np.int16)).cuda()
    img = L.RawImage(md, dev_samples)
    area = img.metadata.active_area
    cam_xyz = synth.camera_to_xyz_for(cam)
    node = L.CameraToWorkingNode(
        L.ExposureNode(L.WhiteBalanceNode(L.RawUnpackNode(img, L.MENON), 1.0, 1.0, 1.0), 0.0),
        L.CameraColorTransform(cam_xyz, L.WorkingSpace.LinearProPhotoD50), area)
    compiled = compile_chain(node)
    out = torch.empty((area.height, area.width, 3), dtype=torch.float32, device="cuda")
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    level = L.RenderLevel()
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        DeviceRender.run(compiled, area, level, out=out, flags=flags, stream=stream)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        DeviceRender.run(compiled, area, level, out=out, flags=flags, stream=stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    assert int(flags.item()) == 0

    # fwd+bwd: the forward render plus the hand-written adjoint of the same
    # chain (dL/d normalised raw over the whole active area + dL/d gains,
    # exposure, matrix) for a synthetic upstream gradient.
    from librawops_amd import _capi as C
    import ctypes
    gout = torch.randn((area.height, area.width, 3), dtype=torch.float32, device="cuda", generator=None)
    graw = torch.zeros((area.height, area.width), dtype=torch.float32, device="cuda")
    gpar = torch.zeros(13, dtype=torch.float64, device="cuda")
    g3 = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    mm = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(compiled_matrix(node), np.float32)])
    sp = ctypes.c_void_p(stream.cuda_stream)

    def backward():
        C.check(C.lib().rawops_render_backward(img._handle, C.DEMOSAIC_MENON, area.c(), g3, 1.0, mm,
                                               ctypes.c_void_p(gout.data_ptr()), area.width,
                                               ctypes.c_void_p(graw.data_ptr()), ctypes.c_void_p(gpar.data_ptr()), sp))
    for _ in range(max(1, args.warmup)):
        backward()
    torch.cuda.synchronize()
    bsteps = max(1, args.steps // 2)
    b0 = [torch.cuda.Event(enable_timing=True) for _ in range(bsteps)]
    b1 = [torch.cuda.Event(enable_timing=True) for _ in range(bsteps)]
    t1 = time.perf_counter()
    for i in range(bsteps):
        DeviceRender.run(compiled, area, level, out=out, flags=flags, stream=stream)
        b0[i].record(stream)
        backward()
        b1[i].record(stream)
    torch.cuda.synchronize()
    fb_wall = time.perf_counter() - t1
    bwd_ms = sum(a.elapsed_time(b) for a, b in zip(b0, b1)) / bsteps

    px = area.width * area.height
    ms_per_step = wall * 1e3 / args.steps
    value = px * args.steps / wall / 1e6
    achieved = BYTES_PER_PX * px / (kernel_ms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline(meta, samples, cam_xyz, threads=min(16, os.cpu_count() or 1))
    line = {
        "metric": "Mpix/s 24MP Menon render (fwd, fwd+bwd) + achieved HBM GB/s vs peak",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded gratings/rectangles/gradients, Poisson-Gaussian noise)",
        "config": {"workload": "Menon render_image, 6048x4024 BGGR phase(1,0) stride 6080, active 6016x3992, "
                               "per-site black, seeded camera->ProPhoto 3x3, FP32 linear out",
                   "pixels": px, "tile": "64x32 fused LDS tile, one launch per frame"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "render_kernel<Menon,64,32>", "kernel_ms": round(kernel_ms, 4),
                     "algorithmic_bytes_per_launch": BYTES_PER_PX * px},
        "cpu_baseline": cpu,
    }
    line["fwd_bwd"] = {"value": round(px * bsteps / fb_wall / 1e6, 2), "unit": "Mpix/s",
                       "ms_per_step": round(fb_wall * 1e3 / bsteps, 4), "backward_kernel_ms": round(bwd_ms, 4),
                       "backward": "menon_backward<64,32>: dL/d raw (24 MP plane) + dL/d gains, exposure, 3x3"}
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
        line["fwd_bwd"]["gpu_fwd_bwd_over_cpu_fwd"] = round(line["fwd_bwd"]["value"] / cpu["value"], 1)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
