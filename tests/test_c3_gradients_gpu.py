"""The guided-filter chain's gradients at its configured size: working-Y
guided filter (Rec.2020 guide) -> +0.7 EV -> 16-point monotone cubic curve on
256x256 crops of the 6000x4000 lognormal raster (mu -2, sigma 1.5, 1 % of
samples above 1), r in {4, 8, 16} x eps in {1e-4, 1e-3, 1e-2}.  The
hand-written adjoints (GuidedFilterFn, ExposureFn, ToneCurveFn) are checked
against central finite differences of the FP64 oracle restatements of the
same chain (SpatialOps.cpp:168-287 without float32 rounding, ToneOps.cpp:
795-844): a directional derivative of the input gradient along a random
direction over the whole crop, per-pixel derivatives at the corners, edges
and interior (evaluated over the pixel's 2r output window), the eps gradient
and all 16 knot-ordinate gradients.  Bar: 1e-4 relative.

The curve is C1 (its second derivative jumps at the knots), so central
differences converge only as O(h): the steps are 1e-7 (relative) for the
whole-crop direction.  The eps gradient sums 65536 per-centre terms of both
signs, and the chain stores the filter output in float32 (as the
reference's tiles do), so the curve's derivative is taken at values within
2^-24 of the FP64 chain's.  Its first reference is therefore the FP64
filter's central difference in eps chained with the exposure and curve
derivatives at the stored float32 values (the derivative of the computation
the device performs); the end-to-end FP64 chain is asserted too wherever
the sum's conditioning (sum of |terms| / |sum|) is below 300, since above
that the float32 storage alone moves the sum by more than 1e-4 (measured:
1.5e-4 at r = 4, eps = 1e-3, conditioning 585)."""
import functools

import numpy as np
import pytest
import torch

from librawops_amd import autograd as A
from librawops_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPACE = 1           # Rec.2020 guide weights
STOPS = 0.7
N = 256


@functools.lru_cache(maxsize=1)
def raster():
    return synth.lognormal_raster(6000, 4000, seed=3)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


@pytest.mark.parametrize("eps", [1e-4, 1e-3, 1e-2])
@pytest.mark.parametrize("radius", [4, 8, 16])
def test_c3_chain_gradients_vs_fd(radius, eps):
    rng = np.random.default_rng(1000 * radius + int(round(-np.log10(eps))))
    R = raster()
    x0, y0 = int(rng.integers(0, 6000 - N)), int(rng.integers(0, 4000 - N))
    p32 = np.ascontiguousarray(R[y0:y0 + N, x0:x0 + N])
    p64 = p32.astype(np.float64)
    knots = synth.monotone_knots(16, seed=3)
    kx = [k[0] for k in knots]
    wgt = rng.normal(size=p32.shape)

    x = torch.from_numpy(p32).cuda().requires_grad_(True)
    e = torch.tensor(eps, device="cuda", requires_grad=True)
    kyt = torch.tensor([k[1] for k in knots], dtype=torch.float32, device="cuda", requires_grad=True)
    q = A.guided_filter(x, e, radius, SPACE)
    y = A.tone_curve(A.exposure(q, torch.tensor(STOPS, device="cuda")), kyt, kx, True)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
    gx = x.grad.double().cpu().numpy()
    ge = float(e.grad)
    gk = kyt.grad.double().cpu().numpy()

    E = float(e.detach().cpu())                      # the float32 eps the device used
    ky = kyt.detach().double().cpu().numpy()          # the float32 ordinates
    m = float(np.float32(A._multiplier(STOPS)))

    def loss(pp, ee=E, yy=ky, vp=None):
        qq = O.guided_filter_f64(pp, radius, ee, SPACE, vp)
        w = wgt if vp is None else wgt[vp[1]:vp[1] + vp[3], vp[0]:vp[0] + vp[2]]
        return float(np.sum(w * O.curve_map_f64(list(zip(kx, yy)), qq * m, True)))

    # input: directional derivative along a random relative direction
    v = rng.normal(size=p64.shape) * p64
    h = 1e-7
    fd = (loss(p64 + h * v) - loss(p64 - h * v)) / (2 * h)
    got = float(np.sum(gx * v))
    assert abs(got - fd) <= 1e-4 * abs(fd), ("direction", got, fd)

    # input: single samples at corners, edges and inside, each over its 2r window
    pts = [(0, 0), (N - 1, 0), (0, N - 1), (N - 1, N - 1), (N // 2, 0), (0, N // 3), (N - 1, N // 2),
           (N // 3, N - 1), (radius, radius + 1), (2 * radius + 1, 2 * radius), (N // 2, N // 2), (97, 181)]
    fds, gots = [], []
    for px, py in pts:
        c = int(rng.integers(0, 3))
        vx0, vy0 = max(0, px - 2 * radius), max(0, py - 2 * radius)
        vp = (vx0, vy0, min(N, px + 2 * radius + 1) - vx0, min(N, py + 2 * radius + 1) - vy0)
        hh = 1e-4 * max(p64[py, px, c], 1e-2)
        pp, pm = p64.copy(), p64.copy()
        pp[py, px, c] += hh
        pm[py, px, c] -= hh
        fds.append((loss(pp, vp=vp) - loss(pm, vp=vp)) / (2 * hh))
        gots.append(gx[py, px, c])
    assert rel(gots, fds) <= 1e-4, ("pixels", gots, fds)

    # eps
    he = 1e-4 * E
    q64 = O.guided_filter_f64(p64, radius, E, SPACE)
    dq = (O.guided_filter_f64(p64, radius, E + he, SPACE) - O.guided_filter_f64(p64, radius, E - he, SPACE)) / (2 * he)
    K = list(zip(kx, ky))
    x32 = (q64.astype(np.float32) * np.float32(m)).astype(np.float64)    # the stored forward values
    dcurve = (O.curve_map_f64(K, x32 + 1e-7, True) - O.curve_map_f64(K, x32 - 1e-7, True)) / 2e-7
    terms = wgt * dcurve * m * dq
    ref = float(np.sum(terms))
    assert abs(ge - ref) <= 1e-4 * abs(ref), ("eps", ge, ref)
    fe = (loss(p64, ee=E + he) - loss(p64, ee=E - he)) / (2 * he)
    cond = float(np.sum(np.abs(terms)) / abs(fe))
    print(f"c3 r={radius} eps={eps}: deps device {ge:.7g}, chain at float32 values {ref:.7g}, FP64 chain {fe:.7g}, "
          f"conditioning {cond:.0f}", flush=True)
    if cond < 300:
        assert abs(ge - fe) <= 1e-4 * abs(fe), ("eps, FP64 chain", ge, fe, cond)

    # curve knot ordinates
    q64 = q64 * m
    fk = []
    for j in range(16):
        d = 1e-6 * (np.arange(16) == j)
        fk.append((np.sum(wgt * O.curve_map_f64(list(zip(kx, ky + d)), q64, True)) -
                   np.sum(wgt * O.curve_map_f64(list(zip(kx, ky - d)), q64, True))) / 2e-6)
    assert rel(gk, fk) <= 1e-4, ("knots", gk, fk)
