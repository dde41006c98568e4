"""The heterogeneous batch at its full workload (64 requests over 8 sources
2000x1500 .. 6048x4024, RGGB / GRBG, viewports 256^2 .. full frame, mip 0/1/2,
bilinear and Menon, FP32 linear + uint8 sRGB preview; synth.c4_workload),
rendered in one captured launch and compared request by request with the CPU
oracle's render of the same chain (Renderer::render_tiles per request,
RawEngine.cpp:1390-1412): FP32 linear output bit for bit; the uint8 preview
(sRGB encode through powf, then round(clamp(v) 255)) within one code, and
equal wherever the oracle's encoded value is not within 1e-5 (the powf
bar) of a rounding boundary."""
import os
import sys

import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd import synth
from librawops_amd.batch import Batch, BatchRequest
from oracle import oracle as O
from tests.helpers import gpu_image, oracle_meta

pytestmark = pytest.mark.gpu


def build(seed=4):
    sources, reqs, cam_w = synth.c4_workload(seed)
    cam = tuple(synth.camera_to_xyz_for(cam_w))
    imgs = [gpu_image(m, torch.from_numpy(s.view(np.int16)).cuda()) for m, s in sources]
    recipe = L.GraphRecipe(red_gain=synth.C4_GAINS[0], green_gain=synth.C4_GAINS[1], blue_gain=synth.C4_GAINS[2],
                           exposure_stops=synth.C4_STOPS,
                           camera_color=L.CameraColorTransform(cam, L.WorkingSpace.LinearProPhotoD50),
                           output_mode=L.OutputMode.SrgbPreview)
    graphs = [{d: L.ImageGraph(img, recipe, demosaic=L.MENON if d == "menon" else L.BILINEAR)
               for d in ("bilinear", "menon")} for img in imgs]
    return sources, reqs, cam, graphs


def batch_requests(reqs, graphs):
    out = []
    for r in reqs:
        x, y, w, h = r["viewport"]
        level = L.RenderLevel(r["mip"], L.RenderQuality.Final if r["mip"] == 0 else L.RenderQuality.Preview)
        out.append(BatchRequest(graphs[r["source"]][r["demosaic"]].output(), L.Rect(x, y, w, h), level, 0,
                                torch.full((h, w, 3), float("nan"), device="cuda"),
                                torch.zeros((h, w, 3), dtype=torch.uint8, device="cuda")))
    return out


@pytest.fixture(scope="module")
def workload():
    """The workload and the oracle's render of every request (once)."""
    sources, reqs, cam, graphs = build()
    threads = min(16, os.cpu_count() or 1)
    want = []
    for r in reqs:
        meta, samples = sources[r["source"]]
        om = oracle_meta(meta)
        kw = dict(demosaic=r["demosaic"], wb=synth.C4_GAINS, exposure_stops=synth.C4_STOPS, camera_to_xyz=cam)
        lin = O.render_raw(om, samples, O.Chain(output="linear", **kw), r["viewport"], r["mip"], threads=threads)
        enc = O.render_raw(om, samples, O.Chain(output="srgb", **kw), r["viewport"], r["mip"], threads=threads)
        want.append((lin, enc))
    return reqs, graphs, want


@pytest.mark.parametrize("shape", [None, 0, 1, 2])
def test_c4_batch_matches_oracle(workload, shape):
    """shape None: the autotuned tile; 0, 1, 2: 64x32, 64x64, 128x32."""
    reqs, graphs, want = workload
    batch_reqs = batch_requests(reqs, graphs)
    batch = Batch(batch_reqs, shape=shape)
    tw, th, _ = batch.shape()
    print(f"c4: tile {tw}x{th} ({'autotuned' if shape is None else 'fixed'})", flush=True)
    for replay in range(2):
        for r in batch_reqs:
            r.out_rgb.fill_(float("nan"))
            r.out_u8.zero_()
        batch.run()
        torch.cuda.synchronize()
        batch.check()
    for k, (r, br, (lin, enc)) in enumerate(zip(reqs, batch_reqs, want)):
        got = br.out_rgb.cpu().numpy()
        assert got.tobytes() == lin.tobytes(), (k, r, tw, th)
        want8 = O.srgb_u8(enc)
        got8 = br.out_u8.cpu().numpy()
        diff = np.abs(got8.astype(np.int32) - want8.astype(np.int32))
        assert diff.max() <= 1, (k, r, tw, th)
        frac = np.clip(enc, 0, 1).astype(np.float64) * 255.0 + 0.5
        near = np.abs(frac - np.rint(frac)) < 255 * 1e-5
        assert not np.any(diff[~near]), (k, r, tw, th)
    print(f"c4: {len(reqs)} requests ok at {tw}x{th}", flush=True)
    sys.stdout.flush()
