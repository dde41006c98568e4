"""The differentiable chain's parameter gradients at its configured size:
Menon -> WB -> 3x3 -> exposure -> guided-filter detail boost (r = 8,
eps = 1e-3) -> sRGB on 512x512 crops of a 4000x3000 RGGB scene (the
config-2 scene generator), through ``full_chain_crops`` (all crops in one
autograd node, as the Adam fit runs them).  Four of the fit's sixteen crops;
directional derivatives in parameter space (each WB gain, exposure, detail
strength, two random directions over the 3x3 matrix) against central finite
differences of the FP64 chain (the demosaic output is the oracle's, bit-equal
to the device's; every later stage in FP64).  Bar: 1e-4 relative."""
import functools
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd import autograd as A
from librawops_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W, H, C = 4000, 3000, 512


@functools.lru_cache(maxsize=1)
def scene():
    s = synth.scene_bayer(W, H, 0, 0, 0, W, (0, 0, W, H), (512,) * 4, 16383, seed=5)
    rng = np.random.default_rng(5)            # the fit's crop sequence (bench.py C5 leg)
    crops = [(int(rng.integers(0, W - C)) // 2 * 2, int(rng.integers(0, H - C)) // 2 * 2, C, C) for _ in range(16)]
    return s, crops


def srgb64(z):
    return np.where(z <= 0.0031308, 12.92 * z, 1.055 * np.maximum(z, 0.0031308) ** (1 / 2.4) - 0.055)


def test_c5_parameter_gradients_vs_fd():
    s, crops = scene()
    crops = [crops[i] for i in (0, 5, 10, 15)]
    md = L.RawMetadata(W, H, W, L.BayerPattern(0), 0, 0, L.Rect(0, 0, 0, 0), (512,) * 4, (16383,) * 4)
    img = L.RawImage(md, torch.from_numpy(s.view(np.int16)).cuda())
    om = O.Bayer(W, H, W, 0, 0, 0, (0, 0, W, H), (512,) * 4, (16383,) * 4)
    M0 = np.asarray(synth.camera_to_prophoto_seeded(5), np.float32) * np.float32(0.9)
    g0 = np.array([1.1, 1.0, 0.95], np.float32)
    st0, sb0 = np.float32(0.15), np.float32(0.3)
    gains = torch.tensor(g0, device="cuda", requires_grad=True)
    stops = torch.tensor(st0, device="cuda", requires_grad=True)
    mat = torch.tensor(M0, device="cuda", requires_grad=True)
    strength = torch.tensor(sb0, device="cuda", requires_grad=True)
    rng = np.random.default_rng(55)
    wgt = rng.normal(size=(len(crops), C, C, 3))
    y = A.full_chain_crops(img, [L.Rect(*c) for c in crops], gains, stops, mat, strength)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()

    demo = [O.render_raw(om, s, O.Chain(demosaic="menon"), c, threads=4).astype(np.float64) for c in crops]
    # the device's exposure multiplier is float32 2^stops (ExposureNode, RawEngine.cpp:465-468)

    def loss_one(i, g, M, stv, sb):
        lin = np.einsum("rc,hwc->hwr", M, demo[i] * np.asarray(g)[None, None, :]) * 2.0 ** stv
        base = O.guided_filter_f64(lin, 8, 1e-3, 0)
        z = np.clip(lin + sb * (lin - base), 0.0, 1.0)
        return float(np.sum(wgt[i] * srgb64(z)))

    pool = ThreadPoolExecutor(max_workers=len(crops))

    def loss(g, M, stv, sb):
        return sum(pool.map(lambda i: loss_one(i, g, M, stv, sb), range(len(crops))))

    G, M64, S, B = g0.astype(np.float64), M0.astype(np.float64), float(st0), float(sb0)
    h = 1e-7
    checks = []
    for c in range(3):
        d = h * (np.arange(3) == c)
        fd = (loss(G + d, M64, S, B) - loss(G - d, M64, S, B)) / (2 * h)
        checks.append((f"gain{c}", float(gains.grad[c]), fd))
    fd = (loss(G, M64, S + h, B) - loss(G, M64, S - h, B)) / (2 * h)
    checks.append(("stops", float(stops.grad), fd))
    fd = (loss(G, M64, S, B + h) - loss(G, M64, S, B - h)) / (2 * h)
    checks.append(("strength", float(strength.grad), fd))
    gm = mat.grad.double().cpu().numpy()
    for k in range(2):
        dm = rng.normal(size=(3, 3))
        fd = (loss(G, M64 + h * dm, S, B) - loss(G, M64 - h * dm, S, B)) / (2 * h)
        checks.append((f"matrix{k}", float(np.sum(gm * dm)), fd))
    pool.shutdown()
    for name, got, fd in checks:
        assert abs(got - fd) <= 1e-4 * abs(fd), (name, got, fd)
