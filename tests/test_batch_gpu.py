"""hipGraph batch: heterogeneous requests (sources, patterns, viewports,
mip 0/1/2, bilinear/Menon) in one launch must reproduce each request's
stand-alone render bit for bit, for every tile shape (64x32, 64x64, 128x32)
and the autotuned one, across graph replays."""
import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd.batch import Batch, BatchRequest
from tests.helpers import random_samples

pytestmark = pytest.mark.gpu

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


def sources():
    out = []
    for i, (w, h, pat, area) in enumerate([(301, 203, 0, (0, 0, 0, 0)), (256, 190, 2, (5, 3, 241, 181)),
                                           (420, 310, 1, (16, 16, 388, 278))]):
        md = L.RawMetadata(w, h, w, L.BayerPattern(pat), i & 1, 0, L.Rect(*area), (510, 512, 511, 513), (16383,) * 4)
        out.append(L.RawImage(md, random_samples(h, w, seed=40 + i, hi=16383)))
    return out


def preview_graph(img, demosaic, linear=False):
    r = L.GraphRecipe(red_gain=1.9, blue_gain=1.4, exposure_stops=0.3,
                      camera_color=L.CameraColorTransform(CAM, L.WorkingSpace.LinearProPhotoD50),
                      output_mode=L.OutputMode.SrgbPreview)
    return L.ImageGraph(img, r, demosaic=demosaic, linear_output=linear)


def requests(imgs):
    rng = np.random.default_rng(3)
    reqs, singles = [], []
    for k in range(14):
        img = imgs[k % 3]
        dem = L.MENON if k % 2 else L.BILINEAR
        mip = k % 3
        g = preview_graph(img, dem)
        lin = preview_graph(img, dem, linear=True)
        level = L.RenderLevel(mip, L.RenderQuality.Final if mip == 0 else L.RenderQuality.Preview)
        b = L.request_bounds(g.source_bounds(), level)
        w = int(rng.integers(1, b.width + 1))
        h = int(rng.integers(1, b.height + 1))
        vp = L.Rect(b.x + int(rng.integers(0, b.width - w + 1)), b.y + int(rng.integers(0, b.height - h + 1)), w, h)
        rgb = torch.full((h, w, 3), float("nan"), device="cuda")
        u8 = torch.zeros((h, w, 3), dtype=torch.uint8, device="cuda")
        reqs.append(BatchRequest(g.output(), vp, level, 0, rgb, u8))
        singles.append((g, lin, vp, level))
    return reqs, singles


SHAPES = {0: (64, 32), 1: (64, 64), 2: (128, 32)}


@pytest.mark.parametrize("shape", [0, 1, 2, None])
def test_batch_matches_single_renders(shape):
    imgs = sources()
    reqs, singles = requests(imgs)
    batch = Batch(reqs, shape=shape)
    for replay in range(3):
        for r in reqs:
            r.out_rgb.fill_(float("nan"))
            r.out_u8.zero_()
        batch.run()
        torch.cuda.synchronize()
        batch.check()
        torch.zeros(1, device="cuda").sum().item()     # eager work between replays
        for r, (g, lin, vp, level) in zip(reqs, singles):
            req = L.RenderRequest(vp, 64, level)
            exp_lin = L.Renderer().render_image(lin, req).rgb
            exp_srgb = L.Renderer().render_image(g, req).rgb
            got = r.out_rgb.cpu().numpy()
            assert got.tobytes() == exp_lin.tobytes(), (replay, vp, level)
            exp_u8 = (np.clip(exp_srgb, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
            assert np.array_equal(r.out_u8.cpu().numpy(), exp_u8), (replay, vp, level)
    lat = batch.latencies_ms()
    assert len(lat) == len(reqs) and all(v >= 0 for v in lat)
    tw, th, _ = batch.shape()
    assert (tw, th) == SHAPES[shape] if shape is not None else (tw, th) in SHAPES.values()
