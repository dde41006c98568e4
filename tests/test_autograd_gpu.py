"""Hand-written adjoints against FP64 finite differences of the oracle
restatements (frozen Menon decisions make the taken branch exactly linear).
Tolerance: 1e-4 relative to the largest gradient magnitude of the check (the
FP32 forward/adjoint accumulate in FP32/FP64 mixes)."""
import ctypes
import math

import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd import autograd as A
from librawops_amd import synth
from oracle import oracle as O
from tests.helpers import gpu_image, menon_adjoint_by_probing, meta_dict, oracle_meta, random_samples

pytestmark = pytest.mark.gpu

CAMW = np.array([[1.05, -0.03, -0.02], [-0.06, 1.02, 0.04], [0.03, 0.09, 0.88]], np.float32)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-30, np.max(np.abs(b))))


def norm_plane(m, s):
    """FP64 normalised samples of the active area (RawEngine.cpp:89-97)."""
    om = oracle_meta(m)
    ax, ay, aw, ah = om.area
    stride = om.row_stride
    codes = s.reshape(-1, stride)[ay:ay + ah, ax:ax + aw].astype(np.float64)
    out = np.zeros_like(codes)
    for y in range(ah):
        for x in range(aw):
            sx, sy = ax + x, ay + y
            site = (((sy & 1) + om.phase_y) & 1) * 2 + (((sx & 1) + om.phase_x) & 1)
            out[y, x] = (codes[y, x] - om.black[site]) / (om.white[site] - om.black[site])
    return out


def bilinear_f64(om, raw):
    """FP64 same-colour mean over an area-relative plane (RawEngine.cpp:316-339)."""
    ax, ay, aw, ah = om.area
    cells = [[0, 1, 1, 2], [2, 1, 1, 0], [1, 0, 2, 1], [1, 2, 0, 1]]

    def col(x, y):
        sx, sy = ax + x, ay + y
        return cells[om.pattern][(((sy & 1) + om.phase_y) & 1) * 2 + (((sx & 1) + om.phase_x) & 1)]
    out = np.zeros((ah, aw, 3))
    for y in range(ah):
        for x in range(aw):
            for c in range(3):
                if col(x, y) == c:
                    out[y, x, c] = raw[y, x]
                    continue
                acc, n = 0.0, 0
                for yy in range(y - 1, y + 2):
                    for xx in range(x - 1, x + 2):
                        if 0 <= yy < ah and 0 <= xx < aw and col(xx, yy) == c:
                            acc += raw[yy, xx]
                            n += 1
                out[y, x, c] = acc / n if n else 0.0
    return out


@pytest.mark.parametrize("demosaic", ["bilinear", "menon"])
@pytest.mark.parametrize("case", [(22, 18, 0, (0, 0, 0, 0), 0, (0, 0)), (29, 23, 32, (3, 2, 24, 19), 1, (1, 0)),
                                  (8, 7, 0, (1, 1, 5, 4), 3, (0, 1))])
def test_render_chain_gradients_vs_fd(demosaic, case):
    w, h, stride, area, pat, ph = case
    m = meta_dict(w, h, stride, pat, ph[0], ph[1], area, (500, 510, 505, 515), (9000,) * 4)
    s = random_samples(h, stride or w, seed=w * 3 + pat, lo=300, hi=9000)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    a = img.metadata.active_area
    vp = L.Rect(a.x, a.y, a.width, a.height)
    gains = torch.tensor([1.7, 1.0, 1.3], device="cuda", requires_grad=True)
    stops = torch.tensor(0.4, device="cuda", requires_grad=True)
    mat = torch.tensor(CAMW, device="cuda", requires_grad=True)
    out = A.render_chain(img, gains, stops, mat, demosaic, vp)
    rng = np.random.default_rng(1)
    wgt = rng.normal(size=out.shape).astype(np.float32)
    loss = (out * torch.from_numpy(wgt).cuda()).sum()
    loss.backward()
    graw, _ = A.render_backward(img, torch.from_numpy(wgt).cuda(), [1.7, 1.0, 1.3], 0.4, CAMW, demosaic, vp)
    torch.cuda.synchronize()

    raw = norm_plane(m, s)
    if demosaic == "menon":
        _, dec = O.menon_f64(om, raw, om.area)
        demo = lambda r: O.menon_f64(om, r, om.area, hz=dec)[0]
    else:
        demo = lambda r: bilinear_f64(om, r)
    mult = 2.0 ** 0.4
    W64 = wgt.astype(np.float64)
    M64 = CAMW.astype(np.float64)

    def loss64(r, g=(1.7, 1.0, 1.3), mm=mult, M=M64):
        d = demo(r) * np.asarray(g)[None, None, :] * mm
        return float(np.sum(W64 * np.einsum("rc,hwc->hwr", M, d)))

    # raw gradient: exact for the linear taken branch
    fd = np.zeros_like(raw)
    eps = 1e-3
    for y in range(raw.shape[0]):
        for x in range(raw.shape[1]):
            rp = raw.copy(); rp[y, x] += eps
            rm = raw.copy(); rm[y, x] -= eps
            fd[y, x] = (loss64(rp) - loss64(rm)) / (2 * eps)
    assert rel(graw.cpu().numpy(), fd) <= 1e-4
    # parameter gradients
    d0 = demo(raw)
    def gvec(c, sign):
        return [v + sign * 1e-4 * (i == c) for i, v in enumerate((1.7, 1.0, 1.3))]
    fg = [(loss64(raw, g=gvec(c, 1)) - loss64(raw, g=gvec(c, -1))) / 2e-4 for c in range(3)]
    assert rel(gains.grad.cpu().numpy(), fg) <= 1e-4
    fs = (loss64(raw, mm=2.0 ** 0.4001) - loss64(raw, mm=2.0 ** 0.3999)) / 2e-4
    assert abs(float(stops.grad) - fs) <= 1e-4 * abs(fs)
    fM = np.zeros((3, 3))
    for r in range(3):
        for c in range(3):
            Mp = M64.copy(); Mp[r, c] += 1e-4
            Mm = M64.copy(); Mm[r, c] -= 1e-4
            fM[r, c] = (loss64(raw, M=Mp) - loss64(raw, M=Mm)) / 2e-4
    assert rel(mat.grad.cpu().numpy(), fM) <= 1e-4
    assert d0.shape[2] == 3


@pytest.mark.parametrize("vp_rel", [None, (37, 29, 201, 87)])
def test_menon_raw_adjoint_large_frame(vp_rel):
    """A 300x150 active area spans interior tiles (forward recompute fully
    inside the area) as well as border tiles; the raw adjoint of every sample
    is checked against the lattice-probed FP64 oracle adjoint, over the whole
    area and over a viewport whose edges cut through interior tiles."""
    w, h, stride, area = 310, 160, 320, (3, 2, 300, 150)
    m = meta_dict(w, h, stride, 2, 0, 1, area, (500, 510, 505, 515), (9000,) * 4)
    s = random_samples(h, stride, seed=11, lo=300, hi=9000)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    vp = (area[0], area[1], area[2], area[3]) if vp_rel is None else \
        (area[0] + vp_rel[0], area[1] + vp_rel[1], vp_rel[2], vp_rel[3])
    rng = np.random.default_rng(3)
    wgt = rng.normal(size=(vp[3], vp[2], 3)).astype(np.float32)
    gains, stops = [1.7, 1.0, 1.3], 0.4
    graw, gp = A.render_backward(img, torch.from_numpy(wgt).cuda(), gains, stops, CAMW, "menon", L.Rect(*vp))
    torch.cuda.synchronize()
    _, dec = O.menon_f64(om, norm_plane(m, s), om.area)
    ref = menon_adjoint_by_probing(om, dec, wgt.astype(np.float64), vp, gains, 2.0 ** 0.4, CAMW.astype(np.float64))
    assert rel(graw.cpu().numpy(), ref) <= 1e-4


@pytest.mark.parametrize("vp_rel", [None, (37, 29, 201, 87)])
def test_recorded_decisions(vp_rel):
    """rawops_render_record: output bits equal rawops_render's and the decision
    plane equals the FP64 oracle's classifier decisions at every non-green site
    of the viewport.  The adjoint that reads the record (no classifier
    recompute on covered interior tiles) equals the recomputing adjoint: the
    same raw-gradient bits (same gather order) and parameter sums to 1e-12."""
    w, h, stride, area = 310, 160, 320, (3, 2, 300, 150)
    m = meta_dict(w, h, stride, 2, 0, 1, area, (500, 510, 505, 515), (9000,) * 4)
    s = random_samples(h, stride, seed=11, lo=300, hi=9000)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    vp = area if vp_rel is None else (area[0] + vp_rel[0], area[1] + vp_rel[1], vp_rel[2], vp_rel[3])
    gains, stops = [1.7, 1.0, 1.3], 0.4
    out_rec, dec = A.render_record(img, gains, stops, CAMW, L.Rect(*vp))
    plan = A.Program.from_ops(A.C.DEMOSAIC_MENON, False, [(A.C.OP_GAINS, gains), (A.C.OP_SCALE, (A._multiplier(stops),)),
                                                          (A.C.OP_MATRIX, [float(v) for v in CAMW.reshape(-1)])], [])
    out = torch.empty_like(out_rec)
    A.C.check(A.C.lib().rawops_render(plan.handle, img._handle, L.Rect(*vp).c(), 0, A.C.QUALITY_FINAL, A._ptr(out),
                                      None, vp[2], None, A._stream()))
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), out_rec.view(torch.int32))
    _, dec_o = O.menon_f64(om, norm_plane(m, s), om.area)
    d = dec.cpu().numpy().view(np.uint32)
    tiles = (vp[2] + 63) // 64
    checked = 0
    for y in range(vp[1] - area[1], vp[1] - area[1] + vp[3]):
        for x in range(vp[0] - area[0], vp[0] - area[0] + vp[2]):
            sx, sy = area[0] + x, area[1] + y
            site = (((sy & 1) + om.phase_y) & 1) * 2 + (((sx & 1) + om.phase_x) & 1)
            if [[0, 1, 1, 2], [2, 1, 1, 0], [1, 0, 2, 1], [1, 2, 0, 1]][om.pattern][site] == 1:
                continue
            rx, ry = sx - vp[0], sy - vp[1]
            bit = (int(d[((ry // 32) * tiles + rx // 64) * 32 + ry % 32]) >> ((rx // 2) % 32)) & 1
            assert bit == dec_o[y, x], (x, y)
            checked += 1
    assert checked >= vp[2] * vp[3] // 2 - vp[3]
    rng = np.random.default_rng(5)
    gout = torch.from_numpy(rng.normal(size=(vp[3], vp[2], 3)).astype(np.float32)).cuda()
    g0, p0 = A.render_backward(img, gout, gains, stops, CAMW, "menon", L.Rect(*vp))
    g1, p1 = A.render_backward(img, gout, gains, stops, CAMW, "menon", L.Rect(*vp), decisions=dec, recorded=L.Rect(*vp))
    torch.cuda.synchronize()
    assert torch.equal(g0.view(torch.int32), g1.view(torch.int32))
    assert rel(p1.cpu().numpy(), p0.cpu().numpy()) <= 1e-12


@pytest.mark.parametrize("radius", [1, 2, 4])
def test_guided_filter_gradients_vs_fd(radius):
    rng = np.random.default_rng(radius)
    p64 = np.exp(rng.normal(-1.0, 0.7, (11, 13, 3)))
    p32 = p64.astype(np.float32)
    p64 = p32.astype(np.float64)
    wgt = rng.normal(size=p32.shape)
    for space in (0, 1):
        x = torch.from_numpy(p32).cuda().requires_grad_(True)
        e = torch.tensor(1e-3, device="cuda", requires_grad=True)
        q = A.guided_filter(x, e, radius, space)
        (q * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
        f = lambda pp, ee=1e-3: float(np.sum(wgt * O.guided_filter_f64(pp, radius, ee, space)))
        fd = np.zeros_like(p64)
        h = 1e-4
        for idx in np.ndindex(p64.shape):
            pp = p64.copy(); pp[idx] += h
            pm = p64.copy(); pm[idx] -= h
            fd[idx] = (f(pp) - f(pm)) / (2 * h)
        assert rel(x.grad.cpu().numpy(), fd) <= 1e-4, space
        fe = (f(p64, 1e-3 * (1 + 1e-4)) - f(p64, 1e-3 * (1 - 1e-4))) / (2e-7)
        assert abs(float(e.grad) - fe) <= 1e-4 * max(1e-12, abs(fe)) + 1e-6, (float(e.grad), fe)
        # forward agrees with the FP64 ordered restatement
        assert rel(q.detach().cpu().numpy(), O.guided_filter_f64(p64, radius, 1e-3, space)) <= 1e-5


@pytest.mark.parametrize("cubic", [True, False])
@pytest.mark.parametrize("nk", [16, 20])
def test_exposure_curve_backward_matches_the_two_adjoints(cubic, nk):
    """rawops_exposure_curve_backward (exposure then curve, one pass) against
    rawops_curve_backward on m x followed by rawops_exposure_backward: dL/dx bit
    for bit, the FP64 knot and stop sums to 1e-12 (their summation order
    differs).  16 knots take the per-lane accumulator form, 20 the shared one;
    values span both extensions and every piece."""
    C = A.C
    knots = synth.monotone_knots(nk, seed=nk)
    X = (ctypes.c_double * (nk + 6))(*([k[0] for k in knots] + [0.0, 1.0] * 3))
    Y = (ctypes.c_double * (nk + 6))(*([k[1] for k in knots] + [0.0, 1.0] * 3))
    plan = A.Program.from_ops(C.DEMOSAIC_BILINEAR, False, [],
                              [(C.OP_CURVES, (), (X, Y, [0, nk, nk + 2, nk + 4, nk + 6], int(cubic)))])
    m = A._multiplier(0.7)
    rng = np.random.default_rng(nk + cubic)
    for shape in ((700, 1000, 3), (5, 7, 3)):
        x = torch.from_numpy(rng.uniform(-0.1, 1.0, shape).astype(np.float32)).cuda()
        gy = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).cuda()
        n = x.numel()
        lin = x * m
        g1, g2, gf = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
        gk, gkf = (torch.zeros(nk, dtype=torch.float64, device="cuda") for _ in range(2))
        gs, gsf = (torch.zeros(1, dtype=torch.float64, device="cuda") for _ in range(2))
        P = lambda t: ctypes.c_void_p(t.data_ptr())
        C.check(C.lib().rawops_curve_backward(plan.handle, P(lin), P(gy), n, P(g1), P(gk), None))
        C.check(C.lib().rawops_exposure_backward(P(x), P(g1), m, P(g2), P(gs), n, None))
        C.check(C.lib().rawops_exposure_curve_backward(plan.handle, P(x), m, P(gy), n, P(gf), P(gkf), P(gsf), None))
        torch.cuda.synchronize()
        assert torch.equal(gf, g2), shape
        assert np.allclose(gkf.cpu().numpy(), gk.cpu().numpy(), rtol=1e-12, atol=1e-12 * float(gk.abs().max())), shape
        assert abs(float(gsf) - float(gs)) <= 1e-12 * abs(float(gs)), shape
    with pytest.raises(C.InvalidArgument):
        C.check(C.lib().rawops_exposure_curve_backward(plan.handle, P(x), m, P(gy), n, P(gf), P(gkf), None, None))


@pytest.mark.parametrize("cubic", [True, False])
def test_tone_curve_gradients_vs_fd(cubic):
    knots = synth.monotone_knots(16, seed=4)
    kx = [k[0] for k in knots]
    ky = np.array([k[1] for k in knots])
    rng = np.random.default_rng(2)
    v = rng.uniform(-0.2, 4.3, (9, 7, 3)).astype(np.float32)
    wgt = rng.normal(size=v.shape)
    x = torch.from_numpy(v).cuda().requires_grad_(True)
    kyt = torch.tensor(ky, dtype=torch.float32, device="cuda", requires_grad=True)
    y = A.tone_curve(x, kyt, kx, cubic)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
    kyf = kyt.detach().double().cpu().numpy()      # the float32 ordinates the forward used
    f = lambda vals, yy: float(np.sum(wgt * O.curve_map_f64(list(zip(kx, yy)), vals, cubic)))
    fk = np.array([(f(v.astype(np.float64), kyf + 1e-6 * (np.arange(16) == j)) -
                    f(v.astype(np.float64), kyf - 1e-6 * (np.arange(16) == j))) / 2e-6 for j in range(16)])
    assert rel(kyt.grad.cpu().numpy(), fk) <= 1e-4
    vv = v.astype(np.float64)
    fx = (O.curve_map_f64(list(zip(kx, kyf)), vv + 1e-5, cubic) - O.curve_map_f64(list(zip(kx, kyf)), vv - 1e-5, cubic)) / 2e-5
    assert rel(x.grad.cpu().numpy(), wgt * fx) <= 1e-4
    assert rel(y.detach().cpu().numpy(), O.curve_map_f64(list(zip(kx, kyf)), vv, cubic)) <= 1e-6


@pytest.mark.parametrize("cubic", [True, False])
def test_tone_curve_gradients_at_knots_and_outside(cubic):
    """Knot-ordinate gradients where the segment search has edge cases: inputs
    exactly on knots (float-representable abscissae), just below and above
    them, below the first and above the last knot, and many values in one
    segment (several samples of a wavefront share knots)."""
    knots = synth.monotone_knots(16, seed=9)
    kx = [float(np.float32(k[0])) for k in knots]
    ky = np.array([k[1] for k in knots])
    xs = np.array(kx, np.float32)
    vals = np.concatenate([xs, np.nextafter(xs, np.float32(-np.inf)), np.nextafter(xs, np.float32(np.inf)),
                           np.float32([-0.5, -1e-3, 1.5, 7.0]),
                           np.linspace(kx[3], kx[4], 64, dtype=np.float32)]).astype(np.float32)
    rng = np.random.default_rng(5)
    v = rng.permutation(np.resize(vals, 3 * 64 * 5)).reshape(64, 5, 3).astype(np.float32)
    wgt = rng.normal(size=v.shape)
    x = torch.from_numpy(v).cuda().requires_grad_(True)
    kyt = torch.tensor(ky, dtype=torch.float32, device="cuda", requires_grad=True)
    y = A.tone_curve(x, kyt, kx, cubic)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
    kyf = kyt.detach().double().cpu().numpy()
    vv = v.astype(np.float64)
    f = lambda yy: float(np.sum(wgt * O.curve_map_f64(list(zip(kx, yy)), vv, cubic)))
    fk = np.array([(f(kyf + 1e-6 * (np.arange(16) == j)) - f(kyf - 1e-6 * (np.arange(16) == j))) / 2e-6
                   for j in range(16)])
    assert rel(kyt.grad.cpu().numpy(), fk) <= 1e-4
    assert rel(y.detach().cpu().numpy(), O.curve_map_f64(list(zip(kx, kyf)), vv, cubic)) <= 1e-6
    assert np.isfinite(x.grad.cpu().numpy()).all()


def test_point_op_gradients():
    rng = np.random.default_rng(3)
    v = rng.uniform(-0.1, 1.2, (5, 6, 3)).astype(np.float32)
    base = rng.uniform(0.0, 1.0, v.shape).astype(np.float32)
    wgt = rng.normal(size=v.shape).astype(np.float32)
    x = torch.from_numpy(v).cuda().requires_grad_(True)
    b = torch.from_numpy(base).cuda().requires_grad_(True)
    s = torch.tensor(0.6, device="cuda", requires_grad=True)
    st = torch.tensor(0.7, device="cuda", requires_grad=True)
    y = A.srgb_encode(A.detail_boost(A.exposure(x, st), b, s))
    (y * torch.from_numpy(wgt).cuda()).sum().backward()

    def f64(xx, bb, ss, stv):
        z = xx * 2.0 ** stv
        z = z + ss * (z - bb)
        l = np.clip(z, 0, 1)
        return float(np.sum(wgt * np.where(l <= 0.0031308, 12.92 * l, 1.055 * l ** (1 / 2.4) - 0.055)))
    X, B = v.astype(np.float64), base.astype(np.float64)
    h = 1e-6
    fs = (f64(X, B, 0.6 + h, 0.7) - f64(X, B, 0.6 - h, 0.7)) / (2 * h)
    fst = (f64(X, B, 0.6, 0.7 + h) - f64(X, B, 0.6, 0.7 - h)) / (2 * h)
    assert abs(float(s.grad) - fs) <= 1e-4 * abs(fs)
    assert abs(float(st.grad) - fst) <= 1e-4 * abs(fst)


def test_full_chain_crops_matches_per_crop_chain():
    """The batched-crops node runs the per-crop kernels in the same order:
    outputs are bit-identical to ``full_chain`` per crop and the parameter
    gradients equal the per-crop sums (FP64 accumulation order aside)."""
    meta, s, cam = synth.config2_menon(seed=6, scale=0.05)
    img = gpu_image(meta, s)
    crops = [L.Rect(40, 30, 48, 40), L.Rect(101, 57, 48, 40), L.Rect(20, 131, 48, 40)]
    args = dict(gains=[1.2, 1.0, 1.1], stops=0.3, mat=CAMW * 0.6, strength=0.5)

    def leaves():
        return (torch.tensor(args["gains"], device="cuda", requires_grad=True),
                torch.tensor(args["stops"], device="cuda", requires_grad=True),
                torch.tensor(args["mat"], device="cuda", requires_grad=True),
                torch.tensor(args["strength"], device="cuda", requires_grad=True))
    rng = np.random.default_rng(2)
    wgt = torch.from_numpy(rng.normal(size=(len(crops), 40, 48, 3)).astype(np.float32)).cuda()
    p1 = leaves()
    y1 = torch.stack([A.full_chain(img, vp, *p1) for vp in crops])
    (y1 * wgt).sum().backward()
    p2 = leaves()
    y2 = A.full_chain_crops(img, crops, *p2)
    (y2 * wgt).sum().backward()
    assert torch.equal(y1, y2)
    for a, b in zip(p1, p2):
        assert rel(b.grad.cpu().numpy(), a.grad.cpu().numpy()) <= 1e-5


def test_full_chain_parameter_gradients_and_fit():
    """Menon -> WB -> 3x3 -> exposure -> detail boost (r=8, eps 1e-3) -> sRGB on
    one 48x40 crop: parameter gradients vs FP64 FD of the composed oracle
    restatements, then a short Adam fit that must reduce the loss."""
    meta, s, cam = synth.config2_menon(seed=5, scale=0.05)
    img = gpu_image(meta, s)
    vp = L.Rect(40, 30, 48, 40)
    om = oracle_meta(meta)
    raw = norm_plane(meta, s)
    _, dec = O.menon_f64(om, raw, om.area)
    ax, ay = om.area[0], om.area[1]
    demo = O.menon_f64(om, raw, (vp.x, vp.y, vp.width, vp.height), hz=dec)[0]
    rng = np.random.default_rng(7)
    wgt = rng.normal(size=(vp.height, vp.width, 3))

    def f64(g, M, stv, strength):
        lin = np.einsum("rc,hwc->hwr", M, demo * np.asarray(g)[None, None, :]) * 2.0 ** stv
        base = O.guided_filter_f64(lin, 8, 1e-3, 0)
        z = np.clip(lin + strength * (lin - base), 0, 1)
        return float(np.sum(wgt * np.where(z <= 0.0031308, 12.92 * z, 1.055 * z ** (1 / 2.4) - 0.055)))

    g0, st0, sb0 = np.array([1.2, 1.0, 1.1]), 0.3, 0.5
    M0 = CAMW.astype(np.float64) * 0.6
    gains = torch.tensor(g0, dtype=torch.float32, device="cuda", requires_grad=True)
    stops = torch.tensor(st0, device="cuda", requires_grad=True)
    mat = torch.tensor(M0, dtype=torch.float32, device="cuda", requires_grad=True)
    strength = torch.tensor(sb0, device="cuda", requires_grad=True)
    y = A.full_chain(img, vp, gains, stops, mat, strength)
    (y * torch.from_numpy(wgt.astype(np.float32)).cuda()).sum().backward()
    h = 1e-5
    fg = [(f64(g0 + h * (np.arange(3) == c), M0, st0, sb0) - f64(g0 - h * (np.arange(3) == c), M0, st0, sb0)) / (2 * h)
          for c in range(3)]
    assert rel(gains.grad.cpu().numpy(), fg) <= 1e-4
    fst = (f64(g0, M0, st0 + h, sb0) - f64(g0, M0, st0 - h, sb0)) / (2 * h)
    fsb = (f64(g0, M0, st0, sb0 + h) - f64(g0, M0, st0, sb0 - h)) / (2 * h)
    assert abs(float(stops.grad) - fst) <= 1e-4 * abs(fst)
    assert abs(float(strength.grad) - fsb) <= 1e-4 * abs(fsb)

    # Adam fit towards a seeded target render
    with torch.no_grad():
        target = A.full_chain(img, vp, torch.tensor([1.5, 1.0, 1.25], device="cuda"), torch.tensor(0.6, device="cuda"),
                              torch.tensor(CAMW * 0.7, device="cuda"), torch.tensor(0.3, device="cuda"))
    params = [torch.tensor(v, dtype=torch.float32, device="cuda", requires_grad=True)
              for v in ([1.0, 1.0, 1.0], 0.0, CAMW * 0.6, 0.0)]
    opt = torch.optim.Adam(params, lr=0.02)
    first = None
    for step in range(60):
        opt.zero_grad()
        loss = ((A.full_chain(img, vp, *params) - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            params[0].clamp_(min=0.05)
        first = float(loss) if first is None else first
    assert float(loss) < 0.2 * first
    assert math.isfinite(float(loss))
