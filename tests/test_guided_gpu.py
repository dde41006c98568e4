"""GPU parity of the working-Y guided filter (+ exposure + 16-point monotone
curve) against the CPU oracle (ordered centred FP64, SpatialOps.cpp:168-287),
the rational impulse KAT and exact-bit special cases.  Tolerance: 1e-5 of
max(1, |value|) (box-moment FP64 accumulation vs ordered sums; values reach
several hundred, where one float32 ulp exceeds 1e-5 absolute)."""
import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def gpu_guided(raster, space, radius, eps, viewport, stops=0.0, knots=None, mip=0, cubic=True):
    H, W = raster.shape[:2]
    img = L.RasterImage(L.RasterMetadata(W, H, 0, L.WorkingSpace(space)), raster)
    n = L.RasterSourceNode(img)
    n = L.WorkingYGuidedFilterNode(n, L.Rect(0, 0, W, H), L.WorkingYGuidedFilterSettings(radius, eps), max_radius=32)
    if stops != 0.0:
        n = L.ExposureNode(n, stops)
    if knots is not None:
        n = L.ExtendedCurvesNode(n, L.ExtendedCurvesSettings(knots[0], knots[1:], L.CurveInterpolation(int(cubic))))
    q = L.RenderQuality.Final if mip == 0 else L.RenderQuality.Preview
    return L.Renderer().render_image(n, L.Rect(0, 0, W, H), L.RenderRequest(L.Rect(*viewport), 256, L.RenderLevel(mip, q))).rgb


def close(a, b, tol=1e-5):
    return np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))


def test_impulse_kat():
    for space in (0, 1):
        img = np.zeros((1, 5, 3), np.float32)
        img[0, 2] = 1
        out = gpu_guided(img, space, 1, 1.0, (0, 0, 5, 1))
        exp = np.array([3 / 22, 2 / 11, 5 / 11, 2 / 11, 3 / 22])
        assert np.max(np.abs(out[0, :, 0] - exp)) <= 1e-7
        assert np.max(np.abs(out[0, :, 1] - exp)) <= 1e-7


def test_constant_and_extreme_exact_bits():
    fmax = np.finfo(np.float32).max
    tiny = np.float32(1.4e-45)
    for space in (0, 1):
        for vals in ([-0.0, 0.0, -0.0], [fmax, fmax, fmax], [-fmax, fmax, -fmax], [tiny, tiny, tiny], [-.5, .25, 2.0]):
            img = np.array(vals, np.float32).reshape(1, 1, 3)
            for radius in (0, 1, 8):
                for eps in (2.0 ** -24, 65536.0):
                    out = gpu_guided(img, space, radius, eps, (0, 0, 1, 1))
                    assert out.tobytes() == img.tobytes(), (vals, radius, eps)


@pytest.mark.parametrize("radius", [1, 4, 8, 16])
@pytest.mark.parametrize("space", [0, 1])
def test_against_oracle(radius, space):
    raster = synth.lognormal_raster(97, 83, seed=radius + 10 * space)
    knots = synth.monotone_knots(16, seed=radius)
    ident = [(0.0, 0.0), (1.0, 1.0)]
    for eps in (1e-4, 1e-3, 1e-2):
        for mip in (0, 1, 2):
            b = O.request_bounds((0, 0, 97, 83), mip)
            for vp in (b, (b[2] // 5, b[3] // 4, b[2] // 2, b[3] // 2)):
                got = gpu_guided(raster, space, radius, eps, vp, stops=0.7, knots=[knots, ident, ident, ident], mip=mip)
                exp = O.guided_chain(raster, space, radius, eps, vp, stops=0.7, knots=[knots, ident, ident, ident],
                                     mip=mip, threads=8)
                err = close(got, exp)
                assert err <= 1e-5, (eps, mip, vp, err)


def test_config3_full_frame_crops():
    """6000x4000 lognormal raster, r = 8, eps 1e-3, +0.7 EV, 16-point curve:
    the full-frame GPU render checked on crops the oracle computes with their
    complete 2r halo."""
    W, H = 6000, 4000
    raster = synth.lognormal_raster(W, H, seed=3)
    knots = synth.monotone_knots(16, seed=3)
    ident = [(0.0, 0.0), (1.0, 1.0)]
    got = gpu_guided(raster, 1, 8, 1e-3, (0, 0, W, H), stops=0.7, knots=[knots, ident, ident, ident])
    rng = np.random.default_rng(0)
    for x, y in [(0, 0), (W - 48, H - 48), (0, H - 48)] + [tuple(rng.integers(0, [W - 48, H - 48])) for _ in range(4)]:
        vp = (int(x), int(y), 48, 48)
        exp = O.guided_chain(raster, 1, 8, 1e-3, vp, stops=0.7, knots=[knots, ident, ident, ident], threads=8)
        err = close(got[y:y + 48, x:x + 48], exp)
        assert err <= 1e-5, (vp, err)


def test_non_finite_input_rejected():
    img = np.ones((4, 4, 3), np.float32)
    dev = torch.from_numpy(img).cuda()
    dev[1, 1, 1] = float("nan")
    raster = L.RasterImage(L.RasterMetadata(4, 4, 0, L.WorkingSpace.LinearProPhotoD50), dev)
    n = L.WorkingYGuidedFilterNode(L.RasterSourceNode(raster), L.Rect(0, 0, 4, 4), L.WorkingYGuidedFilterSettings(1, 1e-3))
    with pytest.raises(L.RawOpsError):
        L.Renderer().render_image(n, L.Rect(0, 0, 4, 4), L.RenderRequest(L.Rect(0, 0, 4, 4)))


@pytest.mark.parametrize("radius", [0, 3, 8])
def test_stack_matches_single_images(radius):
    """rawops_guided_filter_stack / rawops_guided_backward_stack over a stack
    of equal-size images equal the single-image calls on each image (output
    and dL/dp bits; dL/deps to FP64 accumulation order)."""
    from librawops_amd import _capi as C
    from librawops_amd import autograd as A
    n, h, w = 5, 37, 53
    stack = torch.from_numpy(np.stack([synth.lognormal_raster(w, h, seed=70 + i) for i in range(n)])).cuda()
    vp = C.Rect(3, 2, w - 7, h - 5)
    plan = A._empty_plan()
    expose = A.Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [(C.OP_SCALE, (1.5,))])
    for post in (plan, expose):
        want = torch.empty((n, vp.height, vp.width, 3), dtype=torch.float32, device="cuda")
        for i in range(n):
            src = A._Raster(stack[i].contiguous(), 0)
            C.check(C.lib().rawops_guided_filter(src.h, vp, 0, radius, 1e-3, post.handle, A._ptr(want[i]), vp.width,
                                                 None, 0, None, A._stream()))
        got = torch.empty_like(want)
        whole = A._Raster(stack.view(n * h, w, 3), 0)
        C.check(C.lib().rawops_guided_filter_stack(whole.h, n, vp, radius, 1e-3, post.handle, A._ptr(got), vp.width,
                                                   None, A._stream()))
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (radius, post is plan)
    if radius == 0:
        return
    rng = np.random.default_rng(radius)
    gq = torch.from_numpy(rng.normal(size=(n, vp.height, vp.width, 3)).astype(np.float32)).cuda()
    want, got = torch.empty_like(stack), torch.empty_like(stack)
    ge_want = torch.zeros(1, dtype=torch.float64, device="cuda")
    ge_got = torch.zeros(1, dtype=torch.float64, device="cuda")
    for i in range(n):
        src = A._Raster(stack[i].contiguous(), 0)
        C.check(C.lib().rawops_guided_backward(src.h, radius, 1e-3, vp, A._ptr(gq[i]), vp.width, A._ptr(want[i]),
                                               A._ptr(ge_want), None, 0, A._stream()))
    whole = A._Raster(stack.view(n * h, w, 3), 0)
    C.check(C.lib().rawops_guided_backward_stack(whole.h, n, radius, 1e-3, vp, A._ptr(gq), vp.width, A._ptr(got),
                                                 A._ptr(ge_got), A._stream()))
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int32), want.view(torch.int32))
    assert abs(ge_got.item() - ge_want.item()) <= 1e-9 * max(1.0, abs(ge_want.item()))


@pytest.mark.parametrize("cubic", [True, False])
@pytest.mark.parametrize("nk", [2, 5, 16, 17, 40])
def test_staged_curves_against_oracle(cubic, nk):
    """The filter's LDS-staged curves (one record per knot count: linear
    extensions left of the first and right of the last knot, cubic or
    secant segments between) against the oracle's ExtendedCurvesNode
    (ToneOps.cpp:795-844): master and channel curves, non-monotone knots
    (zero and sign-changing tangents), knot counts on both knot searches
    (two levels up to 16, binary above), outputs landing exactly on knots
    and outside the knot range."""
    rng = np.random.default_rng(nk + 100 * int(cubic))
    H, W = 40, 53
    raster = rng.lognormal(-1.0, 1.2, (H, W, 3)).astype(np.float32)
    xs = np.linspace(0.0, 2.0, nk)
    xs[1:-1] += rng.uniform(-0.3, 0.3, nk - 2) * (2.0 / (nk - 1))
    master = [(float(x), float(y)) for x, y in zip(xs, rng.uniform(-0.3, 1.8, xs.size))]
    def chan(seed):
        r = np.random.default_rng(seed)
        cx = np.linspace(-0.5, 2.0, 4 + seed % 3)
        return [(float(x), float(y)) for x, y in zip(cx, np.sort(r.uniform(-0.4, 2.2, cx.size)))]
    knots = [master, chan(1), [(0.0, 0.0), (1.0, 1.0)], chan(2)]
    for vp in [(0, 0, W, H), (5, 3, 31, 22)]:
        got = gpu_guided(raster, 0, 3, 1e-3, vp, stops=0.0, knots=knots, cubic=cubic)
        exp = O.guided_chain(raster, 0, 3, 1e-3, vp, stops=0.0, knots=knots, cubic=cubic)
        assert close(got, exp) <= 1e-5, (vp, close(got, exp))
    # A constant raster filters to itself, so the curves see exactly these
    # values: knots of the master curve, points between, beyond both ends.
    for v in [xs[0], xs[1], xs[-1], 0.5 * (xs[0] + xs[1]), -1.0, 7.5, float(xs[xs.size // 2])]:
        flat = np.full((9, 70, 3), np.float32(v), np.float32)
        got = gpu_guided(flat, 0, 2, 1e-3, (0, 0, 70, 9), knots=knots, cubic=cubic)
        exp = O.guided_chain(flat, 0, 2, 1e-3, (0, 0, 70, 9), knots=knots, cubic=cubic)
        assert close(got, exp) <= 1e-5, (v, close(got, exp))
