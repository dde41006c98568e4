"""The guided-filter chain's gradients at its configured size: working-Y
guided filter (Rec.2020 guide) -> +0.7 EV -> 16-point monotone cubic curve on
256x256 crops of the 6000x4000 lognormal raster (mu -2, sigma 1.5, 1 % of
samples above 1), r in {4, 8, 16} x eps in {1e-4, 1e-3, 1e-2}.  The
hand-written adjoints (GuidedFilterFn, ExposureFn, ToneCurveFn) are checked
against FP64 finite differences of the oracle's guided filter
(SpatialOps.cpp:168-287 without float32 rounding): central differences of
the FP64 filter output along a random direction over the whole crop, along
single samples at the corners, edges and interior (over the sample's 2r
output window), and in eps, chained with the exposure and curve derivatives
(ToneOps.cpp:795-844; the curve's derivative by a 1e-7 central difference of
the FP64 curve at each value), and all 16 knot-ordinate gradients by
central differences of the FP64 curve.  Bar: 1e-4 relative.

Differencing only through the filter keeps the steps off the curve's knots
(the curve is C1: its second derivative jumps there, so differences taken
through it converge only as O(h)).  The chain stores the filter output in
float32, as the reference's tiles do, so the device takes the curve's
derivative at values a few float32 ulps from the FP64 chain's.  Each
derivative is therefore compared with the chain rule at the curve input the
device stored (read back: the derivative of the computation the device
performs) and, where
the sum's conditioning (sum of |terms| / |sum|) is below 300, also with the
chain rule at the FP64 chain's values: above that, the float32 storage alone
moves a sum of terms of both signs by more than 1e-4 (measured: 1.5e-4 for
the eps gradient at r = 4, eps = 1e-3, conditioning 585)."""
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
    xe = A.exposure(q, torch.tensor(STOPS, device="cuda"))
    y = A.tone_curve(xe, kyt, kx, True)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
    gx = x.grad.double().cpu().numpy()
    ge = float(e.grad)
    gk = kyt.grad.double().cpu().numpy()

    E = float(e.detach().cpu())                      # the float32 eps the device used
    ky = kyt.detach().double().cpu().numpy()          # the float32 ordinates
    K = list(zip(kx, ky))
    m = float(np.float32(A._multiplier(STOPS)))

    def gf(pp, ee=E, vp=None):
        return O.guided_filter_f64(pp, radius, ee, SPACE, vp)

    def dcurve(x):
        return (O.curve_map_f64(K, x + 1e-7, True) - O.curve_map_f64(K, x - 1e-7, True)) / 2e-7

    q64 = gf(p64)
    x64 = q64 * m
    x32 = xe.detach().double().cpu().numpy()      # the curve input the forward stored
    up64 = wgt * dcurve(x64) * m          # dL/dq of the FP64 chain
    up32 = wgt * dcurve(x32) * m          # dL/dq at the stored values

    def check(name, got, dq, sl=np.s_[:]):
        t32, t64 = up32[sl] * dq, up64[sl] * dq
        ref32, ref64 = float(np.sum(t32)), float(np.sum(t64))
        cond = float(np.sum(np.abs(t64)) / max(1e-300, abs(ref64)))
        assert abs(got - ref32) <= 1e-4 * abs(ref32), (name, got, ref32, ref64, cond)
        if cond < 300:
            assert abs(got - ref64) <= 1e-4 * abs(ref64), (name, got, ref32, ref64, cond)
        return cond

    # input: a random relative direction over the whole crop
    v = rng.normal(size=p64.shape) * p64
    h = 1e-5
    dq = (gf(p64 + h * v) - gf(p64 - h * v)) / (2 * h)
    cond = check("direction", float(np.sum(gx * v)), dq)

    # input: single samples at corners, edges and inside, each over its 2r window
    pts = [(0, 0), (N - 1, 0), (0, N - 1), (N - 1, N - 1), (N // 2, 0), (0, N // 3), (N - 1, N // 2),
           (N // 3, N - 1), (radius, radius + 1), (2 * radius + 1, 2 * radius), (N // 2, N // 2), (97, 181)]
    got, ref32, ref64 = [], [], []
    for px, py in pts:
        c = int(rng.integers(0, 3))
        vx0, vy0 = max(0, px - 2 * radius), max(0, py - 2 * radius)
        vp = (vx0, vy0, min(N, px + 2 * radius + 1) - vx0, min(N, py + 2 * radius + 1) - vy0)
        hh = 1e-5 * max(p64[py, px, c], 1e-2)
        pp, pm = p64.copy(), p64.copy()
        pp[py, px, c] += hh
        pm[py, px, c] -= hh
        dqw = (gf(pp, vp=vp) - gf(pm, vp=vp)) / (2 * hh)
        sl = np.s_[vp[1]:vp[1] + vp[3], vp[0]:vp[0] + vp[2]]
        got.append(float(gx[py, px, c]))
        ref32.append(float(np.sum(up32[sl] * dqw)))
        ref64.append(float(np.sum(up64[sl] * dqw)))
    # relative to the largest of the twelve (a sample's own gradient can be near zero)
    assert rel(got, ref32) <= 1e-4 and rel(got, ref64) <= 1e-4, ("samples", got, ref32, ref64)

    # eps
    he = 1e-5 * E
    dqe = (gf(p64, ee=E + he) - gf(p64, ee=E - he)) / (2 * he)
    cond_e = check("eps", ge, dqe)
    print(f"c3 r={radius} eps={eps}: conditioning direction {cond:.0f}, eps {cond_e:.0f}", flush=True)

    # curve knot ordinates, at the stored values and at the FP64 chain's
    for x, bar_name in ((x32, "knots"), (x64, "knots, FP64 chain")):
        fk = []
        for j in range(16):
            d = 1e-6 * (np.arange(16) == j)
            fk.append((np.sum(wgt * O.curve_map_f64(list(zip(kx, ky + d)), x, True)) -
                       np.sum(wgt * O.curve_map_f64(list(zip(kx, ky - d)), x, True))) / 2e-6)
        assert rel(gk, fk) <= 1e-4, (bar_name, gk, fk)
