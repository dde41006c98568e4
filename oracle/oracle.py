"""CPU oracle for the tiled RAW render path -- TEST INFRASTRUCTURE ONLY.

Python composition layer over ``liboracle.so`` (the C restatement in
``rawops_oracle.c``).  It strings the restated primitives together in the same
order the reference's node chain does (RawEngine.cpp:1306-1333 for the RAW
graph, RawEngine.cpp:580-601 for mip anchors, SpatialOps.cpp:312-317 and
ToneOps.cpp:833-844 for the raster guided-filter chain).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product path never does.

Colour matrices are restated in pure Python float64 (IEEE double, no FMA), in
the reference's exact operation order (RawEngine.cpp:116-219).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Rect(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class BayerStruct(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("row_stride", ctypes.c_uint32), ("pattern", ctypes.c_uint32),
                ("phase_x", ctypes.c_uint32), ("phase_y", ctypes.c_uint32),
                ("area", Rect),
                ("black", ctypes.c_uint16 * 4), ("white", ctypes.c_uint16 * 4)]


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.ro_demosaic_bilinear.argtypes = [P, P, Rect, P]
        L.ro_demosaic_menon.argtypes = [P, P, Rect, P]
        L.ro_demosaic_menon.restype = ctypes.c_int
        L.ro_demosaic_menon_bands.argtypes = [P, P, Rect, P, ctypes.c_int]
        L.ro_scale3.argtypes = [P, ctypes.c_size_t, P]
        L.ro_scale.argtypes = [P, ctypes.c_size_t, ctypes.c_float]
        L.ro_mat3.argtypes = [P, ctypes.c_size_t, P]
        L.ro_reduce_mip.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_uint32, P]
        L.ro_tone_curve.argtypes = [P, ctypes.c_size_t, ctypes.c_float, ctypes.c_float]
        L.ro_clip.argtypes = [P, ctypes.c_size_t]
        L.ro_srgb_encode.argtypes = [P, ctypes.c_size_t]
        L.ro_srgb_u8.argtypes = [P, ctypes.c_size_t, P]
        L.ro_exp2f.argtypes = [ctypes.c_float]
        L.ro_exp2f.restype = ctypes.c_float
        L.ro_curves.argtypes = [P, ctypes.c_size_t, P, P, P, ctypes.c_int]
        L.ro_guided_filter.argtypes = [P, Rect, Rect, Rect, ctypes.c_uint32, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, P]
        L.ro_guided_filter_bands.argtypes = [P, Rect, Rect, Rect, ctypes.c_uint32, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, P, ctypes.c_int]
        L.ro_local_statistics.argtypes = [P, Rect, Rect, Rect, ctypes.c_uint32, P, P, P]
        L.ro_division_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.ro_division_mismatches.restype = ctypes.c_longlong
        _LIB = L
    return _LIB


def division_mismatches(d_lo=1, d_hi=65535, planted=False):
    """Operand pairs where the kernels' divide-free sample normalisation
    differs from the IEEE float quotient (RawEngine.cpp:89-97)."""
    return int(lib().ro_division_mismatches(d_lo, d_hi, 1 if planted else 0))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ metadata --

@dataclass
class Bayer:
    """Normalised RawMetadata (RawEngine.hpp:57-65; ctor RawEngine.cpp:223-250)."""
    width: int
    height: int
    row_stride: int = 0
    pattern: int = 0          # 0 RGGB, 1 BGGR, 2 GRBG, 3 GBRG
    phase_x: int = 0
    phase_y: int = 0
    area: tuple = (0, 0, 0, 0)
    black: tuple = (0, 0, 0, 0)
    white: tuple = (65535, 65535, 65535, 65535)

    def __post_init__(self):
        if self.row_stride == 0:
            self.row_stride = self.width
        if tuple(self.area) == (0, 0, 0, 0):
            self.area = (0, 0, self.width, self.height)
        self.area = tuple(int(v) for v in self.area)

    def struct(self):
        s = BayerStruct()
        s.width, s.height, s.row_stride = self.width, self.height, self.row_stride
        s.pattern, s.phase_x, s.phase_y = self.pattern, self.phase_x, self.phase_y
        s.area = Rect(*self.area)
        for i in range(4):
            s.black[i] = self.black[i]
            s.white[i] = self.white[i]
        return s


WEIGHTS = {0: (0.28807112822929337, 0.00008565396060525903),   # ProPhoto (ToneOps.cpp:11-13)
           1: (0.26270021201126703, 0.059301716469861945)}     # Rec.2020


# ---------------------------------------------------------------- primitives --

def demosaic(meta: Bayer, samples: np.ndarray, rect, algorithm="bilinear"):
    """RawUnpackNode::render RawEngine.cpp:299-344 / MenonDemosaic.cpp:177-218."""
    L = lib()
    x, y, w, h = rect
    out = np.zeros((h, w, 3), np.float32)
    s = np.ascontiguousarray(samples, np.uint16)
    st = meta.struct()
    aw, ah = meta.area[2], meta.area[3]
    if algorithm == "menon" and aw > 1 and ah > 1:
        if w and h:
            L.ro_demosaic_menon(ctypes.byref(st), _p(s), Rect(x, y, w, h), _p(out))
    elif w and h:
        L.ro_demosaic_bilinear(ctypes.byref(st), _p(s), Rect(x, y, w, h), _p(out))
    return out


def menon_bands(meta: Bayer, samples, rect, threads):
    L = lib()
    x, y, w, h = rect
    out = np.zeros((h, w, 3), np.float32)
    s = np.ascontiguousarray(samples, np.uint16)
    L.ro_demosaic_menon_bands(ctypes.byref(meta.struct()), _p(s), Rect(x, y, w, h), _p(out), threads)
    return out


def scale3(rgb, gains):
    g = np.asarray(gains, np.float32)
    lib().ro_scale3(_p(rgb), rgb.size // 3, _p(g))


def scale(v, k):
    lib().ro_scale(_p(v), v.size, float(np.float32(k)))


def mat3(rgb, m):
    mm = np.asarray(m, np.float32).reshape(9)
    lib().ro_mat3(_p(rgb), rgb.size // 3, _p(mm))


def reduce_mip(native, mip, ow, oh):
    nh, nw = native.shape[:2]
    out = np.zeros((oh, ow, 3), np.float32)
    lib().ro_reduce_mip(_p(np.ascontiguousarray(native)), nw, nh, mip, ow, oh, _p(out))
    return out


def tone_curve(v, shoulder, gamma):
    inv = float(np.float32(1.0) / np.float32(gamma))
    lib().ro_tone_curve(_p(v), v.size, float(np.float32(shoulder)), inv)


def clip(v):
    if lib().ro_clip(_p(v), v.size):
        raise ArithmeticError("non-finite value at output boundary")


def srgb_encode(v):
    if lib().ro_srgb_encode(_p(v), v.size):
        raise ArithmeticError("non-finite value at sRGB output boundary")


def srgb_u8(encoded):
    out = np.zeros(encoded.shape, np.uint8)
    lib().ro_srgb_u8(_p(np.ascontiguousarray(encoded, np.float32)), encoded.size, _p(out))
    return out


def exp2f(stops):
    return np.float32(lib().ro_exp2f(float(np.float32(stops))))


def curves(rgb, knots, cubic):
    """knots: list of 4 lists of (x, y) (master, R, G, B)."""
    xs, ys, off = [], [], [0]
    for k in knots:
        for px, py in k:
            xs.append(px)
            ys.append(py)
        off.append(len(xs))
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    off = np.asarray(off, np.uint32)
    if lib().ro_curves(_p(rgb), rgb.size // 3, _p(xs), _p(ys), _p(off), int(cubic)):
        raise ValueError("extended curves output is not finite")


def guided_filter(inp, needed, output, image, radius, eps, space, threads=1):
    """working_y_guided_filter_rgb SpatialOps.cpp:168-287 on a tile covering
    ``needed`` (output grown by 2r, clipped to image)."""
    wr, wb = WEIGHTS[space]
    out = np.zeros((output[3], output[2], 3), np.float32)
    inp = np.ascontiguousarray(inp, np.float32)
    if threads > 1:
        st = lib().ro_guided_filter_bands(_p(inp), Rect(*needed), Rect(*output), Rect(*image),
                                          radius, eps, wr, wb, _p(out), threads)
    else:
        st = lib().ro_guided_filter(_p(inp), Rect(*needed), Rect(*output), Rect(*image),
                                    radius, eps, wr, wb, _p(out))
    if st:
        raise OverflowError("guided filter invalid intermediate")
    return out


def local_statistics(inp, needed, output, image, radius):
    """local_statistics_rgb SpatialOps.cpp:79-121 on a tile covering ``needed``
    (output grown by radius, clipped to image): (mean, variance, count), each
    (h, w, 3)."""
    inp = np.ascontiguousarray(inp, np.float32)
    assert inp.shape == (needed[3], needed[2], 3)
    shape = (output[3], output[2], 3)
    mean, var, cnt = np.zeros(shape), np.zeros(shape), np.zeros(shape, np.uint32)
    lib().ro_local_statistics(_p(inp), Rect(*needed), Rect(*output), Rect(*image), radius, _p(mean), _p(var),
                              _p(cnt))
    return mean, var, cnt


def grow_rect(r, image, radius):
    """expanded()/expand_rect: grow by radius, clip to image (SpatialOps.cpp:32-40)."""
    x0 = max(image[0], r[0] - radius)
    y0 = max(image[1], r[1] - radius)
    x1 = min(image[0] + image[2], r[0] + r[2] + radius)
    y1 = min(image[1] + image[3], r[1] + r[3] + radius)
    return (x0, y0, x1 - x0, y1 - y0)


# -------------------------------------------------------------- colour math --
# RawEngine.cpp:116-219, restated with the same operation order.

def _mul(a, b):
    out = [0.0] * 9
    for row in range(3):
        for col in range(3):
            for k in range(3):
                out[row * 3 + col] += a[row * 3 + k] * b[k * 3 + col]
    return out


def _mulv(a, v):
    return [a[0] * v[0] + a[1] * v[1] + a[2] * v[2],
            a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
            a[6] * v[0] + a[7] * v[1] + a[8] * v[2]]


def _inv(m):
    det = (m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6])
           + m[2] * (m[3] * m[7] - m[4] * m[6]))
    if not np.isfinite(det) or abs(det) < 1e-12:
        raise ValueError("color matrix is singular")
    d = 1.0 / det
    return [(m[4] * m[8] - m[5] * m[7]) * d, (m[2] * m[7] - m[1] * m[8]) * d,
            (m[1] * m[5] - m[2] * m[4]) * d, (m[5] * m[6] - m[3] * m[8]) * d,
            (m[0] * m[8] - m[2] * m[6]) * d, (m[2] * m[3] - m[0] * m[5]) * d,
            (m[3] * m[7] - m[4] * m[6]) * d, (m[1] * m[6] - m[0] * m[7]) * d,
            (m[0] * m[4] - m[1] * m[3]) * d]


def _white(x, y):
    return [x / y, 1.0, (1.0 - x - y) / y]


def _rgb_to_xyz(xy, wx, wy):
    basis = [0.0] * 9
    for col in range(3):
        x, y = xy[2 * col], xy[2 * col + 1]
        basis[col] = x / y
        basis[3 + col] = 1.0
        basis[6 + col] = (1.0 - x - y) / y
    scales = _mulv(_inv(basis), _white(wx, wy))
    for row in range(3):
        for col in range(3):
            basis[row * 3 + col] *= scales[col]
    return basis


def _d50_to_d65():
    bradford = [0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296]
    d50 = _mulv(bradford, _white(0.3457, 0.3585))
    d65 = _mulv(bradford, _white(0.3127, 0.3290))
    scale_m = [d65[0] / d50[0], 0, 0, 0, d65[1] / d50[1], 0, 0, 0, d65[2] / d50[2]]
    return _mul(_inv(bradford), _mul(scale_m, bradford))


_ROMM = ([0.7347, 0.2653, 0.1596, 0.8404, 0.0366, 0.0001], 0.3457, 0.3585)
_BT2020 = ([0.708, 0.292, 0.170, 0.797, 0.131, 0.046], 0.3127, 0.3290)
_SRGB = ([0.640, 0.330, 0.300, 0.600, 0.150, 0.060], 0.3127, 0.3290)


def xyz_d50_to_working(space):
    if space == 0:
        return _inv(_rgb_to_xyz(*_ROMM))
    return _mul(_inv(_rgb_to_xyz(*_BT2020)), _d50_to_d65())


def working_to_linear_srgb_f32(space):
    s2x = _rgb_to_xyz(*_SRGB)
    if space == 0:
        m = _mul(_inv(s2x), _mul(_d50_to_d65(), _rgb_to_xyz(*_ROMM)))
    else:
        m = _mul(_inv(s2x), _rgb_to_xyz(*_BT2020))
    return np.asarray(m, np.float64).astype(np.float32)


def working_to_working_f32(source, target):
    m = _mul(xyz_d50_to_working(target), _inv(xyz_d50_to_working(source)))
    return np.asarray(m, np.float64).astype(np.float32)


def camera_to_working_f32(camera_to_xyz, space):
    """CameraToWorkingNode ctor RawEngine.cpp:502-530."""
    cam = [float(v) for v in camera_to_xyz]
    _inv(cam)
    composed = _mul(xyz_d50_to_working(space), cam)
    f32max = float(np.finfo(np.float32).max)
    for v in composed:
        if not np.isfinite(v) or abs(v) > f32max:
            raise ValueError("composed color matrix exceeds float32 range")
    return np.asarray(composed, np.float64).astype(np.float32)


# ---------------------------------------------------------------- chains -----

@dataclass
class Chain:
    """The RAW chain of ImageGraph (RawEngine.cpp:1306-1333) with the demosaic
    selectable and an explicit linear-output mode."""
    demosaic: str = "bilinear"
    wb: tuple = (1.0, 1.0, 1.0)
    exposure_stops: float = 0.0
    camera_to_xyz: tuple | None = None
    working_space: int = 0
    output: str = "linear"            # linear | bounded | srgb
    tone_shoulder: float = 0.25
    tone_gamma: float = 1.0


def request_bounds(area, mip):
    """RawEngine.cpp:31-43."""
    if mip == 0:
        return tuple(area)
    s = 1 << mip
    return (0, 0, -(-area[2] // s), -(-area[3] // s))


def render_raw(meta: Bayer, samples, chain: Chain, viewport, mip=0, threads=1):
    """Whole-viewport render of ``chain`` (what Renderer::render_image returns).
    ``threads`` > 1 runs Menon in row bands (the same pixels: each band
    reads its full support, MenonDemosaic.cpp:177-218 blocks are independent)."""
    area = meta.area
    b = request_bounds(area, mip)
    if (viewport[0] < b[0] or viewport[1] < b[1] or viewport[0] + viewport[2] > b[0] + b[2]
            or viewport[1] + viewport[3] > b[1] + b[3]):
        raise IndexError("viewport is outside the active RAW area")   # RawEngine.cpp:22-29

    def native(rect):
        if chain.demosaic == "menon" and threads > 1 and meta.area[2] > 1 and meta.area[3] > 1:
            t = menon_bands(meta, samples, rect, threads)
        else:
            t = demosaic(meta, samples, rect, chain.demosaic)
        scale3(t, chain.wb)
        scale(t, exp2f(chain.exposure_stops))
        if chain.camera_to_xyz is not None:
            mat3(t, camera_to_working_f32(chain.camera_to_xyz, chain.working_space))
        return t

    if mip == 0:
        out = native(viewport)
    else:
        if chain.camera_to_xyz is None:
            raise ValueError("reduced levels need camera calibration")
        s = 1 << mip
        x, y, w, h = viewport
        left = min(x * s, area[2])
        top = min(y * s, area[3])
        right = min((x + w) * s, area[2])
        bottom = min((y + h) * s, area[3])
        mapped = (area[0] + left, area[1] + top, right - left, bottom - top)
        out = reduce_mip(native(mapped), mip, w, h)
    if chain.output == "bounded":
        tone_curve(out, chain.tone_shoulder, chain.tone_gamma)
        clip(out)
    elif chain.output == "srgb":
        mat3(out, working_to_linear_srgb_f32(chain.working_space))
        tone_curve(out, chain.tone_shoulder, chain.tone_gamma)
        srgb_encode(out)
    return out


def reduce_raster(raster, mip, rect):
    """RasterSourceNode::render_level RawEngine.cpp:363-406."""
    if mip == 0:
        x, y, w, h = rect
        return np.ascontiguousarray(raster[y:y + h, x:x + w])
    s = 1 << mip
    H, W = raster.shape[:2]
    x, y, w, h = rect
    x0, y0 = x * s, y * s
    x1, y1 = min((x + w) * s, W), min((y + h) * s, H)
    native = np.ascontiguousarray(raster[y0:y1, x0:x1])
    return reduce_mip(native, mip, w, h)


def guided_chain(raster, space, radius, eps, viewport, stops=0.0, knots=None, cubic=True,
                 mip=0, threads=1):
    """RasterSource -> WorkingYGuidedFilter -> [Exposure] -> [ExtendedCurves]."""
    H, W = raster.shape[:2]
    image = request_bounds((0, 0, W, H), mip)
    needed = grow_rect(viewport, image, 2 * radius)
    inp = reduce_raster(raster, mip, needed)
    out = guided_filter(inp, needed, viewport, image, radius, eps, space, threads)
    if stops != 0.0:
        scale(out, exp2f(stops))
    if knots is not None:
        curves(out, knots, cubic)
    return out


# ------------------------------------------------------- FP64 test entries --

def _lib64():
    L = lib()
    if not getattr(L, "_f64_ready", False):
        P = ctypes.c_void_p
        L.ro_menon_f64.argtypes = [P, P, Rect, P, P, P]
        L.ro_menon_f64.restype = ctypes.c_int
        L.ro_guided_filter_f64.argtypes = [P, Rect, Rect, Rect, ctypes.c_uint32, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, P]
        L.ro_curve_map_f64.argtypes = [P, P, ctypes.c_uint32, ctypes.c_int, P, P, ctypes.c_size_t]
        L._f64_ready = True
    return L


def menon_f64(meta: Bayer, raw, rect, hz=None):
    """Menon on an FP64 normalised-sample plane (area-relative, ah x aw);
    returns (out (h, w, 3) float64, decisions (ah, aw) uint8)."""
    L = _lib64()
    raw = np.ascontiguousarray(raw, np.float64)
    out = np.zeros((rect[3], rect[2], 3), np.float64)
    dec = np.zeros((meta.area[3], meta.area[2]), np.uint8)
    hz_in = np.ascontiguousarray(hz, np.uint8) if hz is not None else None
    L.ro_menon_f64(ctypes.byref(meta.struct()), _p(raw), Rect(*rect),
                   _p(hz_in) if hz_in is not None else None, _p(dec), _p(out))
    return out, dec


def guided_filter_f64(inp, radius, eps, space, viewport=None):
    """FP64 ordered guided filter over a whole (H, W, 3) float64 image."""
    L = _lib64()
    H, W = inp.shape[:2]
    wr, wb = WEIGHTS[space]
    vp = viewport or (0, 0, W, H)
    out = np.zeros((vp[3], vp[2], 3), np.float64)
    L.ro_guided_filter_f64(_p(np.ascontiguousarray(inp, np.float64)), Rect(0, 0, W, H), Rect(*vp),
                           Rect(0, 0, W, H), radius, eps, wr, wb, _p(out))
    return out


def curve_map_f64(knots, values, cubic=True):
    L = _lib64()
    xs = np.asarray([k[0] for k in knots], np.float64)
    ys = np.asarray([k[1] for k in knots], np.float64)
    v = np.ascontiguousarray(values, np.float64)
    out = np.zeros_like(v)
    L.ro_curve_map_f64(_p(xs), _p(ys), len(xs), int(cubic), _p(v), _p(out), v.size)
    return out


def histogram_rgb(rgb, bins, lower, upper):
    """Per-channel histogram of a float32 (h, w, 3) image, restating the
    per-sample loop of histogram_rgb (ImageAnalysis.cpp:52-63): nonfinite
    samples are tallied and skipped; the finite extrema include outliers;
    double(sample) < lower / > upper are tallied as underflow / overflow; else
    bin = min(bins - 1, uint32((double(sample) - lower) / span * bins)).
    Returns [(counts uint64[bins], underflow, overflow, nonfinite, minimum,
    maximum)] for R, G, B (extrema None when no sample is finite)."""
    span = upper - lower
    out = []
    flat = np.asarray(rgb, np.float32).reshape(-1, 3)
    for c in range(3):
        s = flat[:, c]
        fin = np.isfinite(s)
        f = s[fin]
        mn = float(f.min()) if f.size else None
        mx = float(f.max()) if f.size else None
        d = f.astype(np.float64)
        under, over = d < lower, d > upper
        inside = d[~under & ~over]
        b = np.minimum(bins - 1, ((inside - lower) / span * bins).astype(np.uint32))
        counts = np.bincount(b, minlength=bins).astype(np.uint64)
        out.append((counts, int(under.sum()), int(over.sum()), int((~fin).sum()), mn, mx))
    return out
