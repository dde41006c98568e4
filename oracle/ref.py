"""ctypes wrapper over the compiled reference harness (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY.  ``oracle/_ref/libref.so`` is compiled from the
reference's own sources by ``make -C oracle ref`` wherever /root/reference is
present; where it is absent (e.g. a machine that only has this repository) the
tests that need it are skipped and the oracle restatement stands alone.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .oracle import Bayer, Chain, Rect, request_bounds

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ERRORS = {1: ValueError, 2: IndexError, 3: ArithmeticError, 4: MemoryError,
          5: OverflowError, 6: RuntimeError, 7: RuntimeError}


def available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libref.so"))


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(_HERE, "_ref", "libref.so"))
        P = ctypes.c_void_p
        L.refh_last_error.restype = ctypes.c_char_p
        L.refh_render_raw.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int, P, ctypes.c_float,
                                      ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_float, Rect, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, P]
        L.refh_guided_chain.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_uint32, ctypes.c_double, ctypes.c_float,
                                        P, P, P, ctypes.c_int, Rect, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, P]
        L.refh_matrices.argtypes = [P, ctypes.c_int, P, P, P]
        L.refh_time_menon_bands.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_float, P, Rect, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
        L.refh_time_requests.argtypes = [P, P, P, ctypes.c_uint32, P, P, P, P, P, ctypes.c_uint32, P,
                                         ctypes.c_float, P, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
        L.refh_time_guided_crops.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_float, P, ctypes.c_uint32,
                                             ctypes.c_double, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_double)]
        L.refh_exposure_multiplier.argtypes = [ctypes.c_float]
        L.refh_exposure_multiplier.restype = ctypes.c_float
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(code):
    if code:
        raise ERRORS.get(code, RuntimeError)(lib().refh_last_error().decode())


def render_raw(meta: Bayer, samples, chain: Chain, viewport, mip=0, quality=None, tile=256):
    L = lib()
    s = np.ascontiguousarray(samples, np.uint16)
    out = np.zeros((viewport[3], viewport[2], 3), np.float32)
    wb = np.asarray(chain.wb, np.float32)
    cam = np.asarray(chain.camera_to_xyz if chain.camera_to_xyz is not None else [0.0] * 9, np.float64)
    mode = {"linear": 0, "bounded": 1, "srgb": 2}[chain.output]
    if quality is None:
        quality = 1 if mip == 0 else 0
    _check(L.refh_render_raw(ctypes.byref(meta.struct()), _p(s), s.size,
                             int(chain.demosaic == "menon"), _p(wb), float(chain.exposure_stops),
                             int(chain.camera_to_xyz is not None), _p(cam), chain.working_space,
                             mode, float(chain.tone_shoulder), float(chain.tone_gamma),
                             Rect(*viewport), mip, quality, tile, _p(out)))
    return out


def guided_chain(raster, space, radius, eps, viewport, stops=0.0, knots=None, cubic=True,
                 mip=0, tile=256):
    L = lib()
    H, W = raster.shape[:2]
    r = np.ascontiguousarray(raster, np.float32)
    out = np.zeros((viewport[3], viewport[2], 3), np.float32)
    xs = ys = off = None
    if knots is not None:
        xs_l, ys_l, off_l = [], [], [0]
        for k in knots:
            for px, py in k:
                xs_l.append(px)
                ys_l.append(py)
            off_l.append(len(xs_l))
        xs = np.asarray(xs_l, np.float64)
        ys = np.asarray(ys_l, np.float64)
        off = np.asarray(off_l, np.uint32)
    quality = 1 if mip == 0 else 0
    _check(L.refh_guided_chain(_p(r), W, H, space, radius, eps, float(stops),
                               _p(xs), _p(ys), _p(off), int(cubic), Rect(*viewport), mip,
                               quality, tile, _p(out)))
    return out


def time_menon_bands(meta: Bayer, samples, wb, stops, camera_to_xyz, rect, threads, reps):
    """Seconds per round of `threads` concurrent band requests over `rect`
    (Menon [-> WB -> exposure] -> camera-to-ProPhoto; wb None: no WB or
    exposure node), one OpenMP thread each."""
    L = lib()
    s = np.ascontiguousarray(samples, np.uint16)
    w = np.asarray(wb, np.float32) if wb is not None else None
    cam = np.asarray(camera_to_xyz, np.float64)
    out = ctypes.c_double()
    _check(L.refh_time_menon_bands(ctypes.byref(meta.struct()), _p(s), s.size, _p(w), float(stops), _p(cam),
                                   Rect(*rect), threads, reps, ctypes.byref(out)))
    return out.value


def time_requests(sources, requests, wb, stops, camera_to_xyz, threads, reps, output="srgb",
                  tone_shoulder=0.25, tone_gamma=1.0):
    """Seconds per round of serving ``requests`` (dicts with source, demosaic,
    mip, viewport, tile_size) over ``sources`` [(Bayer, samples)] on
    ``threads`` worker threads of the compiled reference, one OpenMP thread
    each (Menon / bilinear -> WB -> exposure -> camera-to-ProPhoto, then the
    SrgbPreview stages for output "srgb")."""
    L = lib()
    metas = (type(sources[0][0].struct()) * len(sources))(*[m.struct() for m, _ in sources])
    arrays = [np.ascontiguousarray(s, np.uint16) for _, s in sources]
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    counts = (ctypes.c_size_t * len(arrays))(*[a.size for a in arrays])
    n = len(requests)
    u32 = lambda v: (ctypes.c_uint32 * n)(*v)
    src = u32([r["source"] for r in requests])
    menon = u32([1 if r["demosaic"] == "menon" else 0 for r in requests])
    mip = u32([r["mip"] for r in requests])
    tile = u32([r["tile_size"] for r in requests])
    vps = (Rect * n)(*[Rect(*r["viewport"]) for r in requests])
    w = np.asarray(wb, np.float32)
    cam = np.asarray(camera_to_xyz, np.float64)
    out = ctypes.c_double()
    _check(L.refh_time_requests(metas, ptrs, counts, len(sources), src, menon, vps, mip, tile, n, _p(w),
                                float(stops), _p(cam), {"linear": 0, "srgb": 2}[output], float(tone_shoulder),
                                float(tone_gamma), threads, reps, ctypes.byref(out)))
    return out.value


def time_guided_crops(meta: Bayer, samples, wb, stops, camera_to_xyz, radius, eps, crops, threads, reps):
    """Seconds per round of the full chain's forward (Menon -> WB -> exposure
    -> camera-to-ProPhoto -> guided filter -> sRGB) over ``crops`` (x, y, w, h)
    of one source, on ``threads`` worker threads of the compiled reference."""
    L = lib()
    s = np.ascontiguousarray(samples, np.uint16)
    w = np.asarray(wb, np.float32)
    cam = np.asarray(camera_to_xyz, np.float64)
    rects = (Rect * len(crops))(*[Rect(*c) for c in crops])
    out = ctypes.c_double()
    _check(L.refh_time_guided_crops(ctypes.byref(meta.struct()), _p(s), s.size, _p(w), float(stops), _p(cam),
                                    int(radius), float(eps), rects, len(crops), threads, reps, ctypes.byref(out)))
    return out.value


def matrices(camera_to_xyz, space):
    L = lib()
    cam = np.asarray(camera_to_xyz, np.float64)
    a = np.zeros(9, np.float32)
    b = np.zeros(9, np.float32)
    c = np.zeros(9, np.float32)
    _check(L.refh_matrices(_p(cam), space, _p(a), _p(b), _p(c)))
    return a, b, c


def exposure_multiplier(stops):
    return np.float32(lib().refh_exposure_multiplier(float(np.float32(stops))))


__all__ = ["available", "render_raw", "guided_chain", "matrices", "exposure_multiplier", "time_menon_bands",
           "Bayer", "Chain", "request_bounds"]


def local_statistics(inp, needed, output, image, radius):
    """The compiled reference's local_statistics_rgb on the same tile."""
    L = lib()
    f = L.refh_local_statistics
    P = ctypes.c_void_p
    f.argtypes = [P, Rect, Rect, Rect, ctypes.c_uint32, P, P, P]
    inp = np.ascontiguousarray(inp, np.float32)
    shape = (output[3], output[2], 3)
    mean, var, cnt = np.zeros(shape), np.zeros(shape), np.zeros(shape, np.uint32)
    _check(f(_p(inp), Rect(*needed), Rect(*output), Rect(*image), radius, _p(mean), _p(var), _p(cnt)))
    return mean, var, cnt
