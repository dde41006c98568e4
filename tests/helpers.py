"""Shared builders: the same chain on the GPU path (librawops_amd, through the
C ABI) and on the CPU oracle (oracle/), over one seeded input."""
from __future__ import annotations

import numpy as np

import librawops_amd as L
from oracle import oracle as O


def meta_dict(width, height, stride=0, pattern=0, px=0, py=0, area=(0, 0, 0, 0), black=(0,) * 4,
              white=(65535,) * 4):
    return dict(width=width, height=height, row_stride_samples=stride, pattern=pattern, cfa_phase_x=px,
                cfa_phase_y=py, active_area=tuple(area), black_levels=tuple(black), white_levels=tuple(white))


def gpu_image(m, samples):
    md = L.RawMetadata(m["width"], m["height"], m["row_stride_samples"], L.BayerPattern(m["pattern"]),
                       m["cfa_phase_x"], m["cfa_phase_y"], L.Rect(*m["active_area"]), tuple(m["black_levels"]),
                       tuple(m["white_levels"]))
    return L.RawImage(md, samples)


def oracle_meta(m):
    return O.Bayer(m["width"], m["height"], m["row_stride_samples"], m["pattern"], m["cfa_phase_x"],
                   m["cfa_phase_y"], tuple(m["active_area"]), tuple(m["black_levels"]), tuple(m["white_levels"]))


def gpu_chain(image, chain: O.Chain, anchor=True):
    """Build the node chain ImageGraph would (RawEngine.cpp:1306-1333) with the
    oracle Chain's parameters."""
    ident = L.MENON if chain.demosaic == "menon" else L.BILINEAR
    n = L.RawUnpackNode(image, ident)
    n = L.WhiteBalanceNode(n, *chain.wb)
    n = L.ExposureNode(n, chain.exposure_stops)
    area = image.metadata.active_area
    if chain.camera_to_xyz is not None:
        n = L.CameraToWorkingNode(n, L.CameraColorTransform(tuple(chain.camera_to_xyz), L.WorkingSpace(chain.working_space)),
                                  area if anchor else L.Rect())
    if chain.output == "bounded":
        n = L.OutputClipNode(L.ToneCurveNode(n, chain.tone_shoulder, chain.tone_gamma))
    elif chain.output == "srgb":
        n = L.SrgbEncodeNode(L.ToneCurveNode(L.WorkingToSrgbNode(n), chain.tone_shoulder, chain.tone_gamma))
    return n


def gpu_render(image, chain: O.Chain, viewport, mip=0):
    node = gpu_chain(image, chain)
    q = L.RenderQuality.Final if mip == 0 else L.RenderQuality.Preview
    req = L.RenderRequest(L.Rect(*viewport), 256, L.RenderLevel(mip, q))
    return L.Renderer().render_image(node, image.metadata.active_area, req).rgb


def random_samples(height, stride, seed, lo=0, hi=20000):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, (height, stride)).astype(np.uint16)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def first_mismatch(a, b):
    ai = np.ascontiguousarray(a, np.float32).view(np.uint32)
    bi = np.ascontiguousarray(b, np.float32).view(np.uint32)
    idx = np.argwhere(ai != bi)
    if idx.size == 0:
        return None
    i = tuple(idx[0])
    return i, float(np.asarray(a)[i]), float(np.asarray(b)[i]), int(idx.shape[0])


def menon_adjoint_by_probing(om, dec, wgt, vp, gains, mult, M):
    """dL/d(normalised raw) of L(r) = sum(wgt * (mult * M @ (gains * menon(r))))
    over the viewport vp = (x, y, w, h) (sensor coordinates), with the Menon
    decisions frozen to dec, using the FP64 oracle forward only.

    With frozen decisions the demosaic is linear and homogeneous in the raw
    samples, and a sample moves outputs within Chebyshev distance 4 only
    (directional green 2, green-site chroma 1, final chroma 1; whole-sample
    reflection at the area edge does not lengthen a read).  Samples 10 apart
    therefore have disjoint 9x9 output windows: one forward per lattice
    offset, over the indicator of that lattice, gives every lattice sample's
    adjoint as the window sum of the per-pixel loss density."""
    ax, ay, aw, ah = om.area
    P, R = 10, 4
    vx, vy, vw, vh = vp
    lin = (np.asarray(M, np.float64) * np.asarray(gains, np.float64)[None, :]) * mult   # out = lin @ demosaic
    out = np.zeros((ah, aw))
    for oy in range(P):
        for ox in range(P):
            d = np.zeros((ah, aw))
            d[oy::P, ox::P] = 1.0
            y = O.menon_f64(om, d, om.area, hz=dec)[0][vy - ay:vy - ay + vh, vx - ax:vx - ax + vw]
            lam = np.zeros((ah + 2 * R, aw + 2 * R))
            lam[R + vy - ay:R + vy - ay + vh, R + vx - ax:R + vx - ax + vw] = np.einsum("hwr,rc,hwc->hw", wgt, lin, y)
            S = np.zeros((ah + 2 * R + 1, aw + 2 * R + 1))
            S[1:, 1:] = lam.cumsum(0).cumsum(1)
            ys = np.arange(oy, ah, P)[:, None]
            xs = np.arange(ox, aw, P)[None, :]
            k = 2 * R + 1
            out[oy::P, ox::P] = S[ys + k, xs + k] - S[ys, xs + k] - S[ys + k, xs] + S[ys, xs]
    return out
