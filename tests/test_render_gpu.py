"""GPU parity of the fused render kernels against the CPU oracle and the
reference's frozen Menon golden file.  Integer unpack, site/phase indexing,
bilinear, Menon (FP64), white balance, exposure, 3x3 and the mip reduction are
required bit-exact; stages that call powf (tone gamma != 1, sRGB encode) are
held to 1e-5 max-abs (values are in [0, 1])."""
import numpy as np
import pytest

import librawops_amd as L
from oracle import oracle as O
from tests.golden import menon_fixture
from tests.helpers import (bits_equal, first_mismatch, gpu_chain, gpu_image, gpu_render, meta_dict, oracle_meta,
                           random_samples)

pytestmark = pytest.mark.gpu

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


@pytest.fixture(scope="module")
def golden():
    return menon_fixture.records()


def test_menon_golden_records_bit_exact(golden):
    """93 frozen upstream fixtures (all patterns/phases, signed, thin fallback):
    full area, ROI and poisoned padding (raw_menon_native_tests.cpp:40-102)."""
    bad = []
    for i, r in enumerate(golden):
        img = gpu_image(r, r["samples"])
        area = r["active_area"]
        out = gpu_render(img, O.Chain(demosaic="menon"), area)
        if not bits_equal(out, r["expected"]):
            bad.append((i, first_mismatch(out, r["expected"])))
            continue
        ax, ay, aw, ah = area
        roi = (ax + aw // 3, ay + ah // 3, max(1, aw // 3), max(1, ah // 3))
        sub = gpu_render(img, O.Chain(demosaic="menon"), roi)
        exp = r["expected"][roi[1] - ay:roi[1] - ay + roi[3], roi[0] - ax:roi[0] - ax + roi[2]]
        if not bits_equal(sub, exp):
            bad.append((i, "roi", first_mismatch(sub, exp)))
        poisoned = r["samples"].reshape(r["height"], r["row_stride_samples"]).copy()
        mask = np.ones_like(poisoned, bool)
        mask[ay:ay + ah, ax:ax + aw] = False
        poisoned[mask] = 0
        out2 = gpu_render(gpu_image(r, poisoned), O.Chain(demosaic="menon"), area)
        if not bits_equal(out2, r["expected"]):
            bad.append((i, "poison", first_mismatch(out2, r["expected"])))
    assert not bad, bad[:5]


CASES = [
    # width, height, stride, area
    (64, 48, 0, (0, 0, 0, 0)),
    (71, 53, 80, (3, 2, 61, 47)),
    (300, 140, 310, (5, 7, 290, 130)),
    (40, 20, 40, (0, 0, 40, 1)),      # singleton row: Menon routes to bilinear
    (33, 37, 35, (1, 1, 2, 35)),      # two-wide active area: folding mirror
]


@pytest.mark.parametrize("demosaic", ["bilinear", "menon"])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("phase", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_linear_chain_bit_exact(demosaic, pattern, phase):
    for ci, (w, h, stride, area) in enumerate(CASES):
        m = meta_dict(w, h, stride, pattern, phase[0], phase[1], area, (510, 512, 511, 513), (16383,) * 4)
        s = random_samples(h, stride or w, seed=ci * 100 + pattern * 10 + phase[0] * 2 + phase[1], hi=17000)
        img = gpu_image(m, s)
        om = oracle_meta(m)
        chain = O.Chain(demosaic=demosaic, wb=(2.0, 1.0, 1.5), exposure_stops=0.5, camera_to_xyz=CAM)
        a = img.metadata.active_area
        for vp in [(a.x, a.y, a.width, a.height),
                   (a.x + a.width // 2, a.y + a.height // 3, max(1, a.width // 3), max(1, a.height // 2))]:
            got = gpu_render(img, chain, vp)
            exp = O.render_raw(om, s, chain, vp)
            assert bits_equal(got, exp), (ci, vp, first_mismatch(got, exp))


@pytest.mark.parametrize("demosaic", ["bilinear", "menon"])
@pytest.mark.parametrize("mip", [1, 2])
def test_mip_reduction_bit_exact(demosaic, mip):
    """CameraToWorkingNode::render_level footprints (RawEngine.cpp:580-601),
    odd active-area origins and ragged right/bottom footprints."""
    m = meta_dict(203, 151, 210, 2, 1, 0, (3, 5, 197, 141), (500, 510, 520, 530), (15000,) * 4)
    s = random_samples(151, 210, seed=7, hi=16000)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    chain = O.Chain(demosaic=demosaic, wb=(1.8, 1.0, 1.3), camera_to_xyz=CAM, working_space=1)
    b = O.request_bounds(om.area, mip)
    for vp in [b, (1, 2, b[2] - 3, b[3] - 4), (b[2] - 5, b[3] - 3, 5, 3)]:
        got = gpu_render(img, chain, vp, mip)
        exp = O.render_raw(om, s, chain, vp, mip)
        assert bits_equal(got, exp), (vp, first_mismatch(got, exp))


@pytest.mark.parametrize("mip", [0, 1, 2])
@pytest.mark.parametrize("gamma", [1.0, 2.2])
def test_srgb_preview_chain(mip, gamma):
    """SrgbPreview output: to-sRGB matrix, tone, sRGB encode (powf: 1e-5)."""
    m = meta_dict(130, 98, 0, 0, 0, 0, (0, 0, 0, 0), (512,) * 4, (16383,) * 4)
    s = random_samples(98, 130, seed=11, hi=16383)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    chain = O.Chain(demosaic="menon", wb=(2.0, 1.0, 1.5), exposure_stops=1.0, camera_to_xyz=CAM,
                    output="srgb", tone_gamma=gamma)
    b = O.request_bounds(om.area, mip)
    got = gpu_render(img, chain, b, mip)
    exp = O.render_raw(om, s, chain, b, mip)
    assert np.max(np.abs(got - exp)) <= 1e-5


def test_bounded_output_and_u8():
    m = meta_dict(96, 64, 0, 1, 0, 1, (0, 0, 0, 0), (256,) * 4, (4095,) * 4)
    s = random_samples(64, 96, seed=5, hi=4095)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    chain = O.Chain(demosaic="bilinear", wb=(1.5, 1.0, 1.2), output="bounded", tone_shoulder=0.3)
    got = gpu_render(img, chain, om.area)
    exp = O.render_raw(om, s, chain, om.area)
    assert bits_equal(got, exp), first_mismatch(got, exp)


def test_config1_bilinear_prophoto_full_frame():
    """1024x768 RGGB 14-bit, black 512 / white 16383, WB (2, 1, 1.5), identity
    camera matrix, ProPhoto FP32 linear output."""
    from librawops_amd import synth
    s = synth.gradient_bayer(1024, 768, seed=1)
    m = meta_dict(1024, 768, 0, 0, 0, 0, (0, 0, 0, 0), (512,) * 4, (16383,) * 4)
    img = gpu_image(m, s)
    chain = O.Chain(demosaic="bilinear", wb=(2.0, 1.0, 1.5), camera_to_xyz=(1, 0, 0, 0, 1, 0, 0, 0, 1))
    got = gpu_render(img, chain, (0, 0, 1024, 768))
    exp = O.render_raw(oracle_meta(m), s, chain, (0, 0, 1024, 768))
    assert bits_equal(got, exp), first_mismatch(got, exp)
    # render_tiles with 256 tiles assembles to the same image
    node = __import__("tests.helpers", fromlist=["gpu_chain"]).gpu_chain(img, chain)
    canvas = np.zeros_like(got)

    def cb(tile):
        b = tile.bounds
        canvas[b.y:b.y + b.height, b.x:b.x + b.width] = tile.rgb

    L.Renderer().render_tiles(node, img.metadata.active_area, L.RenderRequest(L.Rect(0, 0, 1024, 768), 256), cb)
    assert bits_equal(canvas, got)


def test_config2_menon_full_frame_bit_exact():
    """6048x4024 BGGR, phase (1, 0), stride 6080, active inset 16, per-site
    black, seeded camera->ProPhoto matrix: the whole 24 MP frame bit-exact."""
    from librawops_amd import synth
    meta, s, cam = synth.config2_menon(seed=2)
    img = gpu_image(meta, s)
    chain = O.Chain(demosaic="menon", camera_to_xyz=synth.camera_to_xyz_for(cam))
    area = meta["active_area"]
    got = gpu_render(img, chain, area)
    om = oracle_meta(meta)
    nat = O.menon_bands(om, s, area, threads=16)
    O.scale3(nat, (1.0, 1.0, 1.0))
    O.scale(nat, O.exp2f(0.0))
    O.mat3(nat, O.camera_to_working_f32(chain.camera_to_xyz, 0))
    assert bits_equal(got, nat), first_mismatch(got, nat)


def test_non_finite_output_raises_domain_error():
    m = meta_dict(32, 32, 0, 0, 0, 0, (0, 0, 0, 0), (0,) * 4, (100,) * 4)
    s = np.full((32, 32), 100, np.uint16)
    img = gpu_image(m, s)
    huge = (1e30, 0, 0, 0, 1e30, 0, 0, 0, 1e30)
    chain = O.Chain(wb=(65536.0, 65536.0, 65536.0), exposure_stops=32.0, camera_to_xyz=huge, output="bounded")
    with pytest.raises(L.DomainError):
        gpu_render(img, chain, (0, 0, 32, 32))


def test_validation_errors():
    m = meta_dict(16, 16)
    img = gpu_image(m, np.zeros((16, 16), np.uint16))
    node = L.RawUnpackNode(img)
    with pytest.raises(L.OutOfRange):
        L.Renderer().render_image(node, img.metadata.active_area, L.RenderRequest(L.Rect(8, 8, 9, 1)))
    with pytest.raises(L.InvalidArgument):
        L.Renderer().render_image(node, img.metadata.active_area,
                                  L.RenderRequest(L.Rect(0, 0, 4, 4), 256, L.RenderLevel(1, L.RenderQuality.Preview)))
    tok = L.CancellationToken()
    tok.cancel()
    with pytest.raises(L.RenderCancelled):
        L.Renderer().render_tiles(node, img.metadata.active_area, L.RenderRequest(L.Rect(0, 0, 4, 4)),
                                  lambda t: None, cancellation=tok)


@pytest.mark.parametrize("mip", [0, 1, 2])
@pytest.mark.parametrize("source,target", [(0, 1), (1, 0)])
def test_working_space_convert_chain(mip, source, target):
    """CameraToWorking -> WorkingSpaceConvertNode (RawEngine.cpp:603-672) on the
    device: the FP32 working->working matrix after the reduction, bit for bit
    (FP32 products, no contraction) against the oracle's chain followed by
    the same matrix."""
    m = meta_dict(146, 102, 150, 2, 1, 0, (3, 2, 139, 97), (510, 512, 511, 513), (16383,) * 4)
    s = random_samples(102, 150, seed=21 + mip, hi=16383)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    chain = O.Chain(demosaic="menon", wb=(1.8, 1.0, 1.3), exposure_stops=0.4, camera_to_xyz=CAM,
                    working_space=source)
    node = L.WorkingSpaceConvertNode(gpu_chain(img, chain), L.WorkingSpace(target))
    q = L.RenderQuality.Final if mip == 0 else L.RenderQuality.Preview
    b = O.request_bounds(om.area, mip)
    vp = (b[0] + 3, b[1] + 1, b[2] - 5, b[3] - 2)
    got = L.Renderer().render_image(node, img.metadata.active_area,
                                    L.RenderRequest(L.Rect(*vp), 64, L.RenderLevel(mip, q))).rgb
    exp = O.render_raw(om, s, chain, vp, mip)
    O.mat3(exp, O.working_to_working_f32(source, target))
    assert bits_equal(got, exp), first_mismatch(got, exp)
    assert node.output_descriptor() == L.ImageDescriptor.scene_linear(L.WorkingSpace(target))


def test_srgb_encode_power_segment_within_bar():
    """The sRGB encode's power segment (x^(1/2.4) from the hardware log2 and
    exp2, RawEngine.cpp:1264-1275) against the exact power in FP64 over every
    eighth float of (0.0031308, 1], plus the segment ends, the linear
    segment and out-of-range inputs: within 1e-6 absolute (bar 1e-5)."""
    import ctypes
    import torch
    from librawops_amd import _capi as C
    lo = np.float32(0.0031308).view(np.uint32)
    hi = np.float32(1.0).view(np.uint32)
    bits = np.arange(lo, hi + 1, 8, dtype=np.uint32)
    extra = np.array([0.0031308, np.nextafter(np.float32(0.0031308), np.float32(1)), 1.0, 0.5, 0.002, 0.0, -0.0,
                      -1.0, 2.0, 1e-30], np.float32)
    x = np.concatenate([bits.view(np.float32), extra])
    dx = torch.from_numpy(x).cuda()
    dy = torch.empty_like(dx)
    C.check(C.lib().rawops_srgb_encode(ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), x.size, None))
    torch.cuda.synchronize()
    y = dy.cpu().numpy().astype(np.float64)
    lin = np.clip(x.astype(np.float64), 0.0, 1.0)
    exact = np.where(lin <= np.float64(np.float32(0.0031308)), np.float64(np.float32(12.92)) * lin,
                     np.float64(np.float32(1.055)) * lin ** (1.0 / 2.4) - np.float64(np.float32(0.055)))
    assert np.max(np.abs(y - exact)) <= 1e-6
    assert y[-3] == 0.0 and y[-4] == 0.0 and y[-5] == 0.0   # -1, -0 and 0 encode to zero
