"""The streaming Menon kernel (mip-0 Menon renders of viewports that fill
the device: a wavefront per 112-column strip walking down 16-row bands,
register windows and lane shifts, an edge form for the strips that reach the
active area's edge), forced on here for every size, against the CPU oracle
bit for bit (MenonDemosaic.cpp:89-174 and the chain
nodes after it).  Viewports at every row and column parity over all four
CFA patterns reach the four compiled (non-green column, red/blue) forms;
chains reach every op form: none, WB / exposure only, 3x3 only, all three,
and the device op lists (tone + clip, sRGB encode), including viewports
whose bands end mid-group."""
import numpy as np
import pytest

import librawops_amd as L
from oracle import oracle as O
from tests.helpers import bits_equal, first_mismatch, gpu_image, meta_dict, oracle_meta, random_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def streaming_path():
    """Every Menon mip-0 render of this module on the streaming kernel (the
    automatic choice keeps viewports this small on the tile kernel)."""
    from librawops_amd import _capi as C
    C.check(C.lib().rawops_set_menon_path(1, 0))
    yield
    C.check(C.lib().rawops_set_menon_path(0, 0))


def test_paths_render_the_same_bits():
    """Streaming (at several band heights) and tile renders of one viewport."""
    from librawops_amd import _capi as C
    w, h, stride, area = 520, 300, 528, (4, 3, 511, 291)
    m = meta_dict(w, h, stride, 1, 1, 0, area, (510, 512, 511, 513), (16383,) * 4)
    s = random_samples(h, stride, seed=77, hi=16383)
    img = gpu_image(m, s)
    node = chain_node(img, CHAINS["exposure_matrix"])
    req = L.RenderRequest(L.Rect(5, 4, 500, 280), 64)
    outs = []
    try:
        for path, band in ((2, 0), (1, 0), (1, 4), (1, 12), (1, 64)):
            C.check(C.lib().rawops_set_menon_path(path, band))
            outs.append(L.Renderer().render_image(node, img.metadata.active_area, req).rgb)
    finally:
        C.check(C.lib().rawops_set_menon_path(1, 0))
    for o in outs[1:]:
        assert bits_equal(o, outs[0])
    assert C.lib().rawops_set_menon_path(4, 0) == C.INVALID_ARGUMENT
    assert C.lib().rawops_set_menon_path(1, 6) == C.INVALID_ARGUMENT

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


def chain_node(img, c: O.Chain):
    n = L.RawUnpackNode(img, L.MENON)
    if c.wb != (1.0, 1.0, 1.0):
        n = L.WhiteBalanceNode(n, *c.wb)
    if c.exposure_stops != 0.0:
        n = L.ExposureNode(n, c.exposure_stops)
    if c.camera_to_xyz is not None:
        n = L.CameraToWorkingNode(n, L.CameraColorTransform(c.camera_to_xyz, L.WorkingSpace(c.working_space)),
                                  img.metadata.active_area)
    if c.output == "bounded":
        n = L.OutputClipNode(L.ToneCurveNode(n, c.tone_shoulder, c.tone_gamma))
    elif c.output == "srgb":
        n = L.SrgbEncodeNode(L.ToneCurveNode(L.WorkingToSrgbNode(n), c.tone_shoulder, c.tone_gamma))
    return n


CHAINS = {
    "none": O.Chain(demosaic="menon"),
    "wb_exposure": O.Chain(demosaic="menon", wb=(1.8, 1.0, 1.3), exposure_stops=0.4),
    "matrix": O.Chain(demosaic="menon", camera_to_xyz=CAM),
    "wb_exposure_matrix": O.Chain(demosaic="menon", wb=(2.0, 1.0, 1.5), exposure_stops=-0.3, camera_to_xyz=CAM),
    "exposure_matrix": O.Chain(demosaic="menon", exposure_stops=0.7, camera_to_xyz=CAM, working_space=1),
    "tone_clip": O.Chain(demosaic="menon", wb=(1.2, 1.0, 1.7), output="bounded", tone_gamma=1.0),
    "srgb": O.Chain(demosaic="menon", camera_to_xyz=CAM, output="srgb", tone_gamma=2.2),
}


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("chain", list(CHAINS))
def test_stream_interior_matches_oracle(pattern, chain):
    w, h, stride, area = 430, 317, 440, (5, 3, 421, 311)
    m = meta_dict(w, h, stride, pattern, pattern & 1, pattern >> 1, area, (510, 512, 511, 513), (16383,) * 4)
    s = random_samples(h, stride, seed=100 + pattern, hi=16383)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    c = CHAINS[chain]
    node = chain_node(img, c)
    ax, ay, aw, ah = area
    for dx, dy, vw, vh in [(0, 0, aw, ah), (1, 0, aw - 1, ah - 2), (0, 1, aw - 3, ah - 1), (7, 9, 300, 201),
                           (2, 5, 257, 163)]:
        vp = (ax + dx, ay + dy, vw, vh)
        got = L.Renderer().render_image(node, img.metadata.active_area, L.RenderRequest(L.Rect(*vp), 64)).rgb
        exp = O.render_raw(om, s, c, vp)
        if c.output == "srgb" or (c.output == "bounded" and c.tone_gamma != 1.0):
            assert np.max(np.abs(got - exp)) <= 1e-5, (vp, chain)
        else:
            assert bits_equal(got, exp), (vp, chain, first_mismatch(got, exp))


@pytest.mark.parametrize("pattern,phase", [(0, (0, 0)), (1, (1, 0)), (2, (0, 1)), (3, (1, 1))])
@pytest.mark.parametrize("area", [(0, 0, 16, 16), (3, 2, 17, 23), (1, 4, 19, 16), (2, 1, 131, 47), (0, 3, 230, 121)])
def test_stream_edges_match_oracle(pattern, phase, area):
    """The edge form (reflected samples, zero gradients outside the area,
    mirrored rows and columns): whole areas from the smallest the streaming
    kernel takes (16 x 16) up, and viewports touching one edge at a time."""
    ax, ay, aw, ah = area
    w, h = ax + aw + 5, ay + ah + 4
    stride = w + 3
    m = meta_dict(w, h, stride, pattern, phase[0], phase[1], area, (510, 512, 511, 513), (16383,) * 4)
    s = random_samples(h, stride, seed=7 * pattern + aw, hi=16383)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    c = CHAINS["wb_exposure_matrix"]
    node = chain_node(img, c)
    vps = [(ax, ay, aw, ah), (ax, ay + 1, aw // 2 + 1, ah - 2), (ax + aw // 2, ay, aw - aw // 2, ah // 2 + 1),
           (ax + 1, ay + ah // 2, aw - 1, ah - ah // 2), (ax + 2, ay + 3, aw - 4, ah - 6)]
    for vp in vps:
        got = L.Renderer().render_image(node, img.metadata.active_area, L.RenderRequest(L.Rect(*vp), 64)).rgb
        exp = O.render_raw(om, s, c, vp)
        assert bits_equal(got, exp), (vp, first_mismatch(got, exp))


@pytest.mark.parametrize("vp_rel", [(0, 0, None, None), (3, 1, 300, 97), (66, 33, 257, 140)])
def test_stream_recording_matches_tile_recording(vp_rel):
    """rawops_render_record on the streaming path (decision bits OR-ed into a
    cleared record by each strip) writes the same record words and output
    bits as the tile kernel (one record word per tile row)."""
    import torch
    from librawops_amd import autograd as A
    w, h, stride, area = 452, 214, 460, (3, 2, 443, 207)
    m = meta_dict(w, h, stride, 2, 0, 1, area, (500, 510, 505, 515), (9000,) * 4)
    s = random_samples(h, stride, seed=13, lo=300, hi=9000)
    img = gpu_image(m, s)
    ax, ay, aw, ah = area
    dx, dy, vw, vh = vp_rel
    vp = L.Rect(ax + dx, ay + dy, vw or aw - dx, vh or ah - dy)
    cam = np.array([[1.05, -0.03, -0.02], [-0.06, 1.02, 0.04], [0.03, 0.09, 0.88]], np.float32)
    from librawops_amd import _capi as C
    try:
        C.check(C.lib().rawops_set_menon_path(2, 0))
        out_t, dec_t = A.render_record(img, [1.7, 1.0, 1.3], 0.4, cam, vp)
        C.check(C.lib().rawops_set_menon_path(1, 0))
        out_s, dec_s = A.render_record(img, [1.7, 1.0, 1.3], 0.4, cam, vp)
        torch.cuda.synchronize()
    finally:
        C.check(C.lib().rawops_set_menon_path(1, 0))
    assert torch.equal(out_s.view(torch.int32), out_t.view(torch.int32))
    n = ((vp.width + 63) // 64) * ((vp.height + 31) // 32) * 32
    assert torch.equal(dec_s.view(torch.int32)[:n], dec_t.view(torch.int32)[:n])


def test_fp32_form_is_an_approximation_of_the_oracle():
    """rawops_set_menon_path(3): the Menon stages of linear chains in FP32.
    Not the reference's arithmetic, so no parity claim: its classifier
    decisions (recorded like the FP64 form's) agree with the FP64 form's on
    all but a small fraction of the non-green sites, and its output stays
    within FP32 rounding of the oracle except near flipped decisions.
    Chains that are not linear (tone + clip here) keep the FP64 kernel, bit
    for bit."""
    import torch
    from librawops_amd import _capi as C
    from librawops_amd import autograd as A
    w, h, stride, area = 600, 420, 608, (4, 3, 590, 410)
    m = meta_dict(w, h, stride, 1, 1, 0, area, (510, 512, 511, 513), (16383,) * 4)
    s = random_samples(h, stride, seed=31, lo=400, hi=16000)
    img = gpu_image(m, s)
    om = oracle_meta(m)
    ax, ay, aw, ah = area
    vp = L.Rect(ax + 2, ay + 1, aw - 5, ah - 3)
    vpt = (vp.x, vp.y, vp.width, vp.height)
    req = L.RenderRequest(vp, 64)
    c = CHAINS["wb_exposure_matrix"]
    cam = np.array([[1.05, -0.03, -0.02], [-0.06, 1.02, 0.04], [0.03, 0.09, 0.88]], np.float32)
    try:
        C.check(C.lib().rawops_set_menon_path(1, 0))
        out64, dec64 = A.render_record(img, [1.7, 1.0, 1.3], 0.4, cam, vp)
        C.check(C.lib().rawops_set_menon_path(3, 0))
        out32, dec32 = A.render_record(img, [1.7, 1.0, 1.3], 0.4, cam, vp)
        approx = L.Renderer().render_image(chain_node(img, c), img.metadata.active_area, req).rgb
        tone = L.Renderer().render_image(chain_node(img, CHAINS["tone_clip"]), img.metadata.active_area, req).rgb
        torch.cuda.synchronize()
    finally:
        C.check(C.lib().rawops_set_menon_path(1, 0))
    assert bits_equal(tone, O.render_raw(om, s, CHAINS["tone_clip"], vpt))
    # decisions: one bit per non-green site of the viewport
    n = ((vp.width + 63) // 64) * ((vp.height + 31) // 32) * 32
    d64 = dec64.view(torch.int32)[:n].cpu().numpy().view(np.uint8)
    d32 = dec32.view(torch.int32)[:n].cpu().numpy().view(np.uint8)
    flips = int(np.unpackbits(d64 ^ d32).sum())
    assert 0 <= flips < 0.01 * (vp.width * vp.height // 2), flips
    # outputs: within FP32 rounding except near flipped decisions
    for got, ref in ((approx, O.render_raw(om, s, c, vpt)), (out32.cpu().numpy(), out64.cpu().numpy())):
        dev = np.abs(got - ref)
        assert float(np.median(dev)) <= 1e-6
        assert float(np.mean(np.any(dev > 1e-5, axis=2))) < 0.05
