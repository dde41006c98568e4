"""Pins the CPU oracle (oracle/) before it is trusted as the checker:
the reference's frozen Menon golden file, the known-answer values of the
reference's own tests, closed forms, brute force on tiny inputs, and -- where
/root/reference is present -- the compiled reference itself (oracle/_ref)."""
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R
from tests.golden import menon_fixture

needs_ref = pytest.mark.skipif(not R.available(), reason="compiled reference (oracle/_ref) not built here")

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


def _bits(a, b):
    return np.ascontiguousarray(a, np.float32).tobytes() == np.ascontiguousarray(b, np.float32).tobytes()


def test_menon_golden_file():
    """93 fixtures generated upstream with colour-demosaicing 0.2.6: bit-exact."""
    recs = menon_fixture.records()
    assert len(recs) == 93
    for r in recs:
        m = O.Bayer(r["width"], r["height"], r["row_stride_samples"], r["pattern"], r["cfa_phase_x"],
                    r["cfa_phase_y"], r["active_area"], r["black_levels"], r["white_levels"])
        assert _bits(O.demosaic(m, r["samples"], m.area, "menon"), r["expected"])


def test_menon_plant_detected():
    """A dropped classifier weight must break the golden parity: the check is
    sharp enough to catch a plausible restatement slip (planted on the CPU)."""
    r = menon_fixture.records()[5]
    m = O.Bayer(r["width"], r["height"], r["row_stride_samples"], r["pattern"], r["cfa_phase_x"],
                r["cfa_phase_y"], r["active_area"], r["black_levels"], r["white_levels"])
    out = O.demosaic(m, r["samples"], m.area, "menon")
    # bilinear is a different algorithm on the same input: must differ
    assert not _bits(O.demosaic(m, r["samples"], m.area, "bilinear"), r["expected"])
    assert _bits(out, r["expected"])


def test_divide_free_normalisation_is_exact():
    """The kernels normalise samples as q0 + fma(-q0, d, n) y (y = RN(1/d)),
    not with a divide; over every integer numerator 0..65535 and denominator
    1..65535 (all white - black levels) it must round like n / d."""
    assert O.division_mismatches() == 0
    assert O.division_mismatches(1, 64, planted=True) > 0


def test_unpack_kats():
    """core_tests.cpp:134-229: signed values, stride/phase/per-site levels."""
    s = np.full(16, 500, np.uint16)
    s[0] = 0
    s[10] = 2000
    m = O.Bayer(4, 4, 0, 0, 0, 0, (0, 0, 0, 0), (100,) * 4, (1000,) * 4)
    t = O.demosaic(m, s, (0, 0, 4, 4))
    assert t[0, 0, 0] == np.float32(-100.0) / np.float32(900.0)
    assert t[2, 2, 0] == np.float32(1900.0) / np.float32(900.0)
    # phase shift: CFA phase x = 1 on a 2x2 moves the red site
    m2 = O.Bayer(2, 2, 0, 0, 1, 0)
    t2 = O.demosaic(m2, np.array([100, 200, 300, 400], np.uint16), (0, 0, 2, 2))
    assert abs(t2[0, 0, 0] - 200 / 65535) < 1e-7 and abs(t2[0, 0, 1] - 100 / 65535) < 1e-7
    assert abs(t2[0, 0, 2] - 300 / 65535) < 1e-7
    # stride 6, active {1,1,2,2}, phase x 1, per-site levels: all 0.5
    md = O.Bayer(4, 4, 6, 0, 1, 0, (1, 1, 2, 2), (100, 200, 300, 400), (1100, 1200, 1300, 1400))
    smp = np.full(24, 65535, np.uint16)
    for y in (1, 2):
        for x in (1, 2):
            site = (y & 1) * 2 + ((x + 1) & 1)
            smp[y * 6 + x] = md.black[site] + 500
    assert np.all(O.demosaic(md, smp, (1, 1, 2, 2)) == np.float32(0.5))


def test_bilinear_brute_force():
    """Independent float32 restatement of the same-colour 3x3 mean
    (python_raw_preview_tests.py:91-120 logic) on a tiny sensor."""
    rng = np.random.default_rng(3)
    W, H = 9, 7
    s = rng.integers(0, 9000, (H, W)).astype(np.uint16)
    m = O.Bayer(W, H, 0, 2, 1, 1, (1, 1, 7, 5), (100, 200, 300, 400), (9100, 9200, 9300, 9400))
    cells = [[0, 1, 1, 2], [2, 1, 1, 0], [1, 0, 2, 1], [1, 2, 0, 1]]
    site = lambda x, y: (((y & 1) + 1) & 1) * 2 + (((x & 1) + 1) & 1)
    col = lambda x, y: cells[2][site(x, y)]
    norm = lambda x, y: (np.float32(s[y, x]) - np.float32(m.black[site(x, y)])) / np.float32(
        m.white[site(x, y)] - m.black[site(x, y)])
    out = O.demosaic(m, s, (1, 1, 7, 5))
    for y in range(1, 6):
        for x in range(1, 8):
            for c in range(3):
                if col(x, y) == c:
                    exp = norm(x, y)
                else:
                    acc, n = np.float32(0), 0
                    for yy in range(y - 1, y + 2):
                        for xx in range(x - 1, x + 2):
                            if 1 <= yy < 6 and 1 <= xx < 8 and col(xx, yy) == c:
                                acc = np.float32(acc + norm(xx, yy))
                                n += 1
                    exp = np.float32(acc / np.float32(n)) if n else np.float32(0)
                assert out[y - 1, x - 1, c] == exp


def test_matrix_kats():
    """ProPhoto / Rec.2020 / sRGB matrix KATs (core_tests.cpp:255-266, 437-441)."""
    pp = O.camera_to_working_f32((1, 0, 0, 0, 1, 0, 0, 0, 1), 0).reshape(3, 3)
    v = pp.astype(np.float64) @ np.array([0.5, 0.25, 0.1])
    assert np.allclose(v, [0.60399, 0.1068, 0.12123], atol=1e-3)
    rec = O.camera_to_working_f32((1, 0, 0, 0, 1, 0, 0, 0, 1), 1).reshape(3, 3)
    assert np.allclose(rec.astype(np.float64) @ np.array([0.5, 0.25, 0.1]), [0.7016773, 0.0718773, 0.1244141], atol=1e-3)
    srgb = O.working_to_linear_srgb_f32(1).reshape(3, 3)
    assert np.allclose(srgb[:, 0], [1.660491, -0.124550, -0.018151], atol=1e-3)
    # neutral preserved by working -> sRGB
    assert np.allclose(O.working_to_linear_srgb_f32(0).reshape(3, 3).sum(1), 1.0, atol=2e-5)
    with pytest.raises(ValueError):
        O.camera_to_working_f32((1, 0, 0, 1, 0, 0, 0, 0, 1), 0)


def test_srgb_kats():
    v = np.array([0.0031308, 0.18, 1.2, -1.0], np.float32)
    O.srgb_encode(v)
    assert abs(v[0] - 0.04045) < 1e-5 and abs(v[1] - 0.461356) < 1e-5 and abs(v[2] - 1.0) < 1e-6 and v[3] == 0.0
    with pytest.raises(ArithmeticError):
        O.srgb_encode(np.array([np.nan], np.float32))
    u = O.srgb_u8(np.array([0.0, 1.0, 0.5, 0.998], np.float32))
    assert list(u) == [0, 255, 128, 254]


def test_mip_reduction_brute_force():
    rng = np.random.default_rng(4)
    nat = rng.normal(size=(11, 13, 3)).astype(np.float32)
    for mip in (1, 2):
        s = 1 << mip
        ow, oh = -(-13 // s), -(-11 // s)
        out = O.reduce_mip(nat, mip, ow, oh)
        for y in range(oh):
            for x in range(ow):
                blk = nat[y * s:(y + 1) * s, x * s:(x + 1) * s].reshape(-1, 3).astype(np.float64)
                exp = np.array([sum(F(float(v)) for v in blk[:, c]) / len(blk) for c in range(3)])
                assert np.all(np.abs(out[y, x].astype(np.float64) - exp.astype(np.float64)) <= 1e-7 * np.abs(exp) + 1e-30)


def test_curves_truth():
    """curves_levels_tests.cpp:14-30: composition/cast truth and flat segments."""
    rgb = np.array([.25, .25, .25, .5029296875, .5029296875, .5029296875], np.float32)
    master = [(0, 0), (.5, .25), (1, 1)]
    ch = [(0, 0), (.5, .75), (1, 1)]
    O.curves(rgb, [master, ch, ch, ch], cubic=True)
    assert rgb[0] == np.float32(.2109375) and rgb[3] == np.float32(.4401211142539978)
    flat = np.array([1.25, 1.5, 1.75], np.float32)
    ident = [(0, 0), (1, 1)]
    O.curves(flat, [[(0, 0), (1, 1), (2, 1), (3, 2)], ident, ident, ident], cubic=True)
    assert list(flat) == [1, 1, 1]


def _rational_gf(rgb, radius, eps, wr, wb):
    """Exact expanded-kernel guided filter on a tiny 1-D/2-D image
    (python_guided_filter_tests.py:104-125 formula), in rationals."""
    H, W = rgb.shape[:2]
    P = [[[F(float(rgb[y, x, c])) for c in range(3)] for x in range(W)] for y in range(H)]
    wr, wb = F(wr), F(wb)
    guide = [[wr * p[0] + (1 - wr - wb) * p[1] + wb * p[2] for p in row] for row in P]

    def window(x, y):
        return [(xx, yy) for yy in range(max(0, y - radius), min(H, y + radius + 1))
                for xx in range(max(0, x - radius), min(W, x + radius + 1))]

    mean = {}
    den = {}
    for y in range(H):
        for x in range(W):
            ids = window(x, y)
            mean[(x, y)] = sum(guide[j][i] for i, j in ids) / len(ids)
            den[(x, y)] = sum((guide[j][i] - mean[(x, y)]) ** 2 for i, j in ids) / len(ids) + F(eps)
    out = np.zeros((H, W, 3), np.float64)
    for y in range(H):
        for x in range(W):
            centers = window(x, y)
            acc = [F(0)] * 3
            for k in centers:
                wk = window(*k)
                for (i, j) in wk:
                    wgt = (1 + (guide[y][x] - mean[k]) * (guide[j][i] - mean[k]) / den[k]) / (len(centers) * len(wk))
                    for c in range(3):
                        acc[c] += wgt * P[j][i][c]
            out[y, x] = [float(a) for a in acc]
    return out


def test_guided_filter_impulse_kat():
    """guided_filter_tests.cpp:46-49: neutral impulse, r=1, eps=1 ->
    [3/22, 2/11, 5/11, 2/11, 3/22]."""
    for space in (0, 1):
        img = np.zeros((1, 5, 3), np.float32)
        img[0, 2] = 1
        out = O.guided_chain(img, space, 1, 1.0, (0, 0, 5, 1))
        exp = [3 / 22, 2 / 11, 5 / 11, 2 / 11, 3 / 22]
        for x in range(5):
            assert np.all(out[0, x] == np.float32(exp[x]))


def test_guided_filter_rational_kernel():
    rng = np.random.default_rng(9)
    img = rng.lognormal(-1.0, 0.8, (5, 6, 3)).astype(np.float32)
    for space in (0, 1):
        wr, wb = O.WEIGHTS[space]
        for radius, eps in ((1, 1e-3), (2, 1e-2)):
            out = O.guided_chain(img, space, radius, eps, (0, 0, 6, 5))
            exp = _rational_gf(img, radius, eps, wr, wb)
            assert np.max(np.abs(out - exp)) <= 1e-6 * max(1.0, np.max(np.abs(exp)))


def _box_gf(rgb, radius, eps, wr, wb):
    """Guided filter by its textbook closed form (He et al.; the expanded
    kernel of python_guided_filter_tests.py:104-125 summed over centres):
    per centre k, a_k = cov_k(g, p) / (var_k(g) + eps), b_k = mean_k(p) -
    a_k mean_k(g) over the window clipped to the image; output = mean over
    the centres whose window holds the pixel of a_k g + b_k.  Box sums by
    2-D prefix sums in float64, independent of the oracle's ordered
    difference sums."""
    p = np.asarray(rgb, np.float64)
    H, W = p.shape[:2]
    g = wr * p[..., 0] + (1.0 - wr - wb) * p[..., 1] + wb * p[..., 2]

    def box(a):
        S = np.zeros((H + 1, W + 1) + a.shape[2:])
        S[1:, 1:] = a.cumsum(0).cumsum(1)
        y0 = np.clip(np.arange(H) - radius, 0, H)[:, None]
        y1 = np.clip(np.arange(H) + radius + 1, 0, H)[:, None]
        x0 = np.clip(np.arange(W) - radius, 0, W)[None, :]
        x1 = np.clip(np.arange(W) + radius + 1, 0, W)[None, :]
        return S[y1, x1] - S[y0, x1] - S[y1, x0] + S[y0, x0]

    n = box(np.ones((H, W)))
    mg = box(g) / n
    mp = box(p) / n[..., None]
    var = box(g * g) / n - mg * mg
    cov = box(g[..., None] * p) / n[..., None] - mg[..., None] * mp
    a = cov / (var + eps)[..., None]
    b = mp - a * mg[..., None]
    return (box(a) * g[..., None] + box(b)) / n[..., None]


@pytest.mark.parametrize("radius", [12, 16])
def test_guided_filter_large_radius_closed_form(radius):
    """Radii the reference does not admit (it stops at 8): the oracle's FP64
    and FP32 filters against the closed form on images wider and taller than
    the 33-pixel window, so windows clip on one side, both sides and neither;
    a planted radius off by one must fail."""
    from librawops_amd import synth as S
    img = S.lognormal_raster(41, 37, seed=radius)
    for space in (0, 1):
        wr, wb = O.WEIGHTS[space]
        for eps in (1e-4, 1e-2):
            want = _box_gf(img, radius, eps, wr, wb)
            got64 = O.guided_filter_f64(img.astype(np.float64), radius, eps, space)
            assert np.max(np.abs(got64 - want)) <= 1e-9 * max(1.0, np.max(np.abs(want)))
            got32 = O.guided_chain(img, space, radius, eps, (0, 0, 41, 37))
            assert np.max(np.abs(got32 - want)) <= 2e-6 * max(1.0, np.max(np.abs(want)))
            wrong = _box_gf(img, radius - 1, eps, wr, wb)
            assert np.max(np.abs(got64 - wrong)) > 1e-4


def test_menon_bands_equal_whole_render():
    """The oracle's banded Menon (the CPU baseline's and the batch test's
    threaded form) renders the same pixels as the one-call form."""
    rng = np.random.default_rng(4)
    s = rng.integers(0, 16383, (131, 214)).astype(np.uint16)
    m = O.Bayer(210, 131, 214, 1, 1, 0, (3, 2, 203, 125), (510, 512, 511, 513), (16383,) * 4)
    for rect in [(3, 2, 203, 125), (20, 9, 97, 61)]:
        whole = O.demosaic(m, s, rect, "menon")
        for threads in (2, 3, 7):
            assert _bits(O.menon_bands(m, s, rect, threads), whole), (rect, threads)


@needs_ref
@pytest.mark.parametrize("demosaic", ["bilinear", "menon"])
def test_oracle_matches_compiled_reference_chains(demosaic):
    rng = np.random.default_rng(12)
    for (w, h, stride, area, pat, ph) in [(70, 52, 72, (3, 2, 61, 47), 1, (1, 0)),
                                          (129, 67, 129, (0, 0, 129, 67), 3, (0, 1)),
                                          (40, 3, 40, (0, 0, 40, 3), 2, (1, 1))]:
        s = rng.integers(0, 17000, (h, stride)).astype(np.uint16)
        m = O.Bayer(w, h, stride, pat, ph[0], ph[1], area, (510, 512, 511, 513), (16383,) * 4)
        for chain in [O.Chain(demosaic=demosaic, wb=(2.0, 1.0, 1.5), exposure_stops=0.7, camera_to_xyz=CAM),
                      O.Chain(demosaic=demosaic, wb=(1.2, 1.0, 1.7), output="bounded", tone_gamma=2.2),
                      O.Chain(demosaic=demosaic, camera_to_xyz=CAM, working_space=1, output="srgb")]:
            for mip in ((0, 1, 2) if chain.camera_to_xyz is not None and chain.output != "bounded" else (0,)):
                b = O.request_bounds(m.area, mip)
                if chain.output == "linear" and mip:
                    continue
                for vp in [b, (b[0] + b[2] // 4, b[1] + b[3] // 4, max(1, b[2] // 2), max(1, b[3] // 2))]:
                    a = O.render_raw(m, s, chain, vp, mip)
                    r = R.render_raw(m, s, chain, vp, mip, tile=37)
                    if chain.output == "linear" or (chain.output == "bounded" and chain.tone_gamma == 1.0):
                        assert _bits(a, r), (chain, mip, vp)
                    else:
                        assert np.max(np.abs(a - r)) <= 1e-6


@needs_ref
def test_compiled_reference_request_pool_times_a_batch():
    """The CPU baseline of the batch leg: requests of two sources served by a
    worker pool of the compiled reference (bad requests raise)."""
    rng = np.random.default_rng(5)
    srcs = []
    for w, h, pat in [(96, 64, 0), (130, 90, 2)]:
        srcs.append((O.Bayer(w, h, w, pat, 0, 0, (0, 0, w, h), (512,) * 4, (16383,) * 4),
                     rng.integers(0, 16383, (h, w)).astype(np.uint16)))
    reqs = [dict(source=0, demosaic="menon", mip=0, viewport=(0, 0, 96, 64), tile_size=32),
            dict(source=1, demosaic="bilinear", mip=1, viewport=(3, 2, 40, 30), tile_size=128),
            dict(source=1, demosaic="menon", mip=2, viewport=(0, 0, 33, 23), tile_size=256)]
    assert R.time_requests(srcs, reqs, (2.0, 1.0, 1.5), 0.3, CAM, 3, 2) > 0.0
    bad = dict(source=0, demosaic="menon", mip=0, viewport=(90, 0, 20, 10), tile_size=32)
    with pytest.raises(IndexError):
        R.time_requests(srcs, [bad], (2.0, 1.0, 1.5), 0.3, CAM, 2, 1)


@needs_ref
def test_compiled_reference_crop_pool_times_the_full_chain():
    """The CPU baseline of the differentiable-chain leg: crops of one source
    through Menon ... guided filter ... sRGB on a worker pool."""
    rng = np.random.default_rng(6)
    m = O.Bayer(120, 80, 120, 0, 0, 0, (0, 0, 120, 80), (512,) * 4, (16383,) * 4)
    s = rng.integers(0, 16383, (80, 120)).astype(np.uint16)
    crops = [(0, 0, 40, 40), (60, 20, 40, 40), (78, 38, 42, 42)]
    assert R.time_guided_crops(m, s, (1.5, 1.0, 1.25), 0.4, CAM, 4, 1e-3, crops, 2, 1) > 0.0
    with pytest.raises(IndexError):
        R.time_guided_crops(m, s, (1.5, 1.0, 1.25), 0.4, CAM, 4, 1e-3, [(100, 0, 40, 40)], 2, 1)


@needs_ref
def test_oracle_matches_compiled_reference_matrices_and_exposure():
    for space in (0, 1):
        a, b, c = R.matrices(CAM, space)
        assert _bits(a, O.camera_to_working_f32(CAM, space))
        assert _bits(b, O.working_to_linear_srgb_f32(space))
        assert _bits(c, O.working_to_working_f32(space, 1 - space))
    for stops in (-3.5, 0.0, 0.7, 1.0, 31.9):
        assert R.exposure_multiplier(stops) == O.exp2f(stops)


@needs_ref
@pytest.mark.parametrize("radius", [1, 4, 8])
def test_oracle_matches_compiled_reference_guided_chain(radius):
    from librawops_amd import synth as S
    img = S.lognormal_raster(37, 29, seed=radius)
    knots = S.monotone_knots(16, seed=radius)
    ident = [(0.0, 0.0), (1.0, 1.0)]
    for space in (0, 1):
        for eps in (1e-4, 1e-2):
            for mip in (0, 1):
                b = O.request_bounds((0, 0, 37, 29), mip)
                vp = (1, 2, b[2] - 2, b[3] - 3)
                a = O.guided_chain(img, space, radius, eps, vp, stops=0.7, knots=[knots, ident, ident, ident], mip=mip)
                r = R.guided_chain(img, space, radius, eps, vp, stops=0.7, knots=[knots, ident, ident, ident], mip=mip)
                assert _bits(a, r), (space, eps, mip)
