"""Seeded synthetic inputs for the render-path workloads (benchmarks, tests).

Each generator returns host numpy data plus the metadata the workload names.
Inputs are produced once on the host and uploaded; they are data, not part of
any timed region.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


def xyz_d50_to_working(space: int) -> np.ndarray:
    out = (ctypes.c_double * 9)()
    C.check(C.lib().rawops_xyz_d50_to_working_matrix(space, out))
    return np.asarray(list(out), np.float64).reshape(3, 3)


def camera_to_prophoto_seeded(seed: int) -> np.ndarray:
    """Seeded camera->ProPhoto matrix with rows summing to 1."""
    rng = np.random.default_rng(seed)
    m = np.eye(3) + rng.uniform(-0.2, 0.2, (3, 3))
    for i in range(3):
        off = m[i].sum() - m[i, i]
        m[i, i] = 1.0 - off
    return m


def camera_to_xyz_for(camera_to_working: np.ndarray, space: int = 0) -> tuple:
    """camera->XYZ(D50) such that the node's composed matrix is camera_to_working."""
    xyz_to_w = xyz_d50_to_working(space)
    return tuple((np.linalg.inv(xyz_to_w) @ camera_to_working).reshape(-1).tolist())


# Colour of each pattern-relative site (0 R, 1 G, 2 B), RawEngine.cpp:79-87.
_CELLS = np.array([[0, 1, 1, 2], [2, 1, 1, 0], [1, 0, 2, 1], [1, 2, 0, 1]])


def cfa_colors(width, height, pattern, phase_x, phase_y):
    ys = (np.arange(height)[:, None] & 1)
    xs = (np.arange(width)[None, :] & 1)
    site = ((ys + phase_y) & 1) * 2 + ((xs + phase_x) & 1)
    return _CELLS[pattern][site]


def gradient_bayer(width=1024, height=768, pattern=0, black=512, white=16383, sigma=20.0, seed=1):
    """Smooth horizontal + vertical gradients plus Gaussian noise, tight rows."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, width)[None, :]
    y = np.linspace(0.0, 1.0, height)[:, None]
    colors = cfa_colors(width, height, pattern, 0, 0)
    scene = np.stack([0.15 + 0.7 * x + 0.1 * y + 0 * y, 0.2 + 0.35 * x + 0.4 * y, 0.7 - 0.5 * x + 0.2 * y])
    level = np.take_along_axis(scene.reshape(3, -1) if False else np.broadcast_to(scene, (3, height, width)),
                               colors[None], 0)[0]
    dn = black + level * (white - black) + rng.normal(0.0, sigma, (height, width))
    return np.clip(np.rint(dn), 0, white).astype(np.uint16)


def scene_rgb(width, height, seed):
    """Gratings (2-32 px periods, random angles), sharp rectangles and smooth
    colour gradients, linear RGB in [0, 1]."""
    rng = np.random.default_rng(seed)
    x = np.arange(width, dtype=np.float32)[None, :]
    y = np.arange(height, dtype=np.float32)[:, None]
    img = np.empty((3, height, width), np.float32)
    img[0] = 0.15 + 0.55 * (x / width) + 0.1 * (y / height)
    img[1] = 0.25 + 0.25 * (x / width) + 0.35 * (y / height)
    img[2] = 0.6 - 0.4 * (x / width) + 0.2 * (y / height)
    for _ in range(max(4, width * height // 1_000_000)):
        w = int(rng.integers(min(64, width), max(65, min(width, 900))))
        h = int(rng.integers(min(64, height), max(65, min(height, 900))))
        x0 = int(rng.integers(0, max(1, width - w)))
        y0 = int(rng.integers(0, max(1, height - h)))
        period = float(rng.uniform(2.0, 32.0))
        theta = float(rng.uniform(0.0, np.pi))
        xx = np.arange(x0, x0 + w, dtype=np.float32)[None, :]
        yy = np.arange(y0, y0 + h, dtype=np.float32)[:, None]
        g = 0.5 + 0.5 * np.sin((2 * np.pi / period) * (xx * np.cos(theta) + yy * np.sin(theta)))
        tint = rng.uniform(0.2, 1.0, 3).astype(np.float32)
        img[:, y0:y0 + h, x0:x0 + w] = (0.1 + 0.8 * g)[None] * tint[:, None, None]
    for _ in range(max(8, width * height // 500_000)):
        w = int(rng.integers(8, max(9, width // 6)))
        h = int(rng.integers(8, max(9, height // 6)))
        x0 = int(rng.integers(0, max(1, width - w)))
        y0 = int(rng.integers(0, max(1, height - h)))
        img[:, y0:y0 + h, x0:x0 + w] = rng.uniform(0.02, 0.98, 3).astype(np.float32)[:, None, None]
    return img


def scene_bayer(width, height, pattern, phase_x, phase_y, stride, area, black, white,
                gain=0.5, read_noise=3.0, seed=2, poison=65535):
    """Poisson-Gaussian noisy mosaic of scene_rgb over the active area; padding
    and inactive photosites are filled with ``poison``."""
    ax, ay, aw, ah = area
    rng = np.random.default_rng(seed + 1000)
    rgb = scene_rgb(aw, ah, seed)
    colors = cfa_colors(width, height, pattern, phase_x, phase_y)[ay:ay + ah, ax:ax + aw]
    level = np.take_along_axis(rgb, colors[None], 0)[0]
    black_site = np.asarray(black, np.float64)
    ys = (np.arange(ay, ay + ah)[:, None] & 1)
    xs = (np.arange(ax, ax + aw)[None, :] & 1)
    site = ((ys + phase_y) & 1) * 2 + ((xs + phase_x) & 1)
    blk = black_site[site]
    span = (white - blk) * 0.92
    electrons = rng.poisson(np.maximum(level * span, 0.0) / gain)
    dn = blk + electrons * gain + rng.normal(0.0, read_noise, (ah, aw))
    out = np.full((height, stride), poison, np.uint16)
    out[ay:ay + ah, ax:ax + aw] = np.clip(np.rint(dn), 0, 65535).astype(np.uint16)
    return out


def config2_menon(seed=2, scale=1.0):
    """6048x4024 BGGR, phase (1, 0), stride 6080, active inset 16, per-site
    black (510, 512, 511, 513), white 16383; ``scale`` < 1 shrinks it for tests."""
    W, H = int(6048 * scale) // 2 * 2, int(4024 * scale) // 2 * 2
    stride = W + 32
    area = (16, 16, W - 32, H - 32)
    meta = dict(width=W, height=H, row_stride_samples=stride, pattern=1, cfa_phase_x=1, cfa_phase_y=0,
                active_area=area, black_levels=(510, 512, 511, 513), white_levels=(16383,) * 4)
    samples = scene_bayer(W, H, 1, 1, 0, stride, area, (510, 512, 511, 513), 16383, seed=seed)
    cam = camera_to_prophoto_seeded(seed)
    return meta, samples, cam


def lognormal_raster(width, height, seed=3, mu=-2.0, sigma=1.5, over=0.01):
    """Scene-linear RGB: lognormal values below 1, with a fraction ``over``
    of samples raised above 1."""
    rng = np.random.default_rng(seed)
    v = np.exp(rng.normal(mu, sigma, (height, width, 3))).astype(np.float32)
    v = np.minimum(v, np.float32(0.999))
    mask = rng.random((height, width, 3)) < over
    v[mask] = (1.0 + np.exp(rng.normal(-1.0, 1.0, int(mask.sum())))).astype(np.float32)
    return v


def monotone_knots(n=16, seed=3):
    """A seeded strictly increasing 16-point tone curve over [0, 4]."""
    rng = np.random.default_rng(seed)
    xs = np.concatenate([[0.0], np.sort(rng.uniform(0.02, 3.9, n - 2)), [4.0]])
    ys = np.cumsum(np.concatenate([[0.0], rng.uniform(0.05, 1.0, n - 1)]))
    ys = ys / ys[-1] * 1.2
    return list(zip(xs.tolist(), ys.tolist()))


# Recipe of every C4 request: WB (1.9, 1, 1.4), +0.3 EV, camera->ProPhoto,
# SrgbPreview output (tone shoulder 0.25, gamma 1).
C4_GAINS = (1.9, 1.0, 1.4)
C4_STOPS = 0.3


def c4_workload(seed=4, count=64):
    """The heterogeneous batch: ``count`` render requests over 8 sources.

    Sources: 14-bit Bayer mosaics of the config-2 scene generator, sized
    2000x1500 .. 6048x4024 (evenly spaced), RGGB for even and GRBG for odd
    source numbers, tight rows, whole-sensor active area, black 512, white
    16383.  Requests: a random source, bilinear (even k) or Menon (odd k),
    level (mip 0, Final) / (mip 1 or 2, Preview), a viewport of log-uniform
    area between 256^2 (at the level) and the whole level bounds at the
    source's aspect ratio and a random offset, and a render_tiles tile size of
    128 / 256 / 512.  Returns (sources [(meta dict, samples)], requests
    [dict(source, demosaic, mip, viewport (x, y, w, h), tile_size)],
    camera->ProPhoto matrix)."""
    rng = np.random.default_rng(seed)
    sizes = [(int(w) // 2 * 2, int(h) // 2 * 2)
             for w, h in zip(np.linspace(2000, 6048, 8), np.linspace(1500, 4024, 8))]
    sources = []
    for i, (w, h) in enumerate(sizes):
        pat = 0 if i % 2 == 0 else 2
        meta = dict(width=w, height=h, row_stride_samples=w, pattern=pat, cfa_phase_x=0, cfa_phase_y=0,
                    active_area=(0, 0, w, h), black_levels=(512,) * 4, white_levels=(16383,) * 4)
        sources.append((meta, scene_bayer(w, h, pat, 0, 0, w, (0, 0, w, h), (512,) * 4, 16383,
                                          seed=seed * 100 + i)))
    requests = []
    for k in range(count):
        i = int(rng.integers(0, 8))
        mip = int(rng.integers(0, 3))
        aw, ah = sizes[i]
        s = 1 << mip
        bw, bh = -(-aw // s), -(-ah // s)          # request_bounds, RawEngine.cpp:31-43
        lo = max(1, 256 >> mip)
        area = float(np.exp(rng.uniform(np.log(lo * lo), np.log(bw * bh))))
        w = int(min(bw, max(lo, round((area * bw / bh) ** 0.5))))
        h = int(min(bh, max(lo, round(area / w))))
        x = int(rng.integers(0, bw - w + 1))
        y = int(rng.integers(0, bh - h + 1))
        requests.append(dict(source=i, demosaic="menon" if k % 2 else "bilinear", mip=mip, viewport=(x, y, w, h),
                             tile_size=(128, 256, 512)[k % 3]))
    return sources, requests, camera_to_prophoto_seeded(seed)
