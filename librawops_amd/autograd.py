"""Differentiable forms of the render path on PyTorch-ROCm.

Each ``torch.autograd.Function`` runs its forward and its hand-written
adjoint as gfx950 kernels through the C ABI, on torch's current stream, with
tensors already in HBM:

* ``render_chain``  RAW -> (bilinear | Menon) -> white balance -> exposure ->
  camera matrix (the reference's node order, RawEngine.cpp:1306-1316); the
  Menon adjoint is the transpose of the branch its classifier takes.
* ``guided_filter``  working-Y guided filter; backward w.r.t. the input pixels
  (through the input-dependent guide) and eps, with the box statistics
  recomputed rather than stored.
* ``exposure``, ``srgb_encode``, ``detail_boost``, ``tone_curve`` (master
  curve, gradients w.r.t. knot ordinates).

There is no CPU execution path.  Non-finite checks of the node API are not run
here (no flag word is passed).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _capi as C
from .graph import Program, RawImage, Rect

_LN2 = math.log(2.0)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _multiplier(stops: float) -> float:
    m = ctypes.c_float()
    C.check(C.lib().rawops_exposure_multiplier(float(np.float32(stops)), ctypes.byref(m)))
    return m.value


class _Raster:
    """Device raster view of a contiguous (H, W, 3) float32 CUDA tensor."""

    def __init__(self, t, space=0):
        assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 3 and t.shape[2] == 3
        self.t = t
        self.h = ctypes.c_void_p()
        self._free = C.lib().rawops_source_destroy   # bound now: usable during interpreter exit
        H, W = t.shape[:2]
        C.check(C.lib().rawops_source_wrap_raster_device(W, H, 0, space, _ptr(t), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._free(self.h)


_EMPTY = None


def _empty_plan():
    global _EMPTY
    if _EMPTY is None:
        _EMPTY = Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [])
    return _EMPTY


def _apply_program(x, program, space=0):
    """Run a reduced-op program over a (H, W, 3) tensor (radius-0 raster pass)."""
    H, W = x.shape[:2]
    src = _Raster(x.contiguous(), space)
    y = torch.empty_like(x)
    C.check(C.lib().rawops_guided_filter(src.h, C.Rect(0, 0, W, H), 0, 0, 1.0, program.handle, _ptr(y), W,
                                         None, 0, None, _stream()))
    return y


# -------------------------------------------------------------- render chain --

class RenderChainFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gains, stops, matrix, image: RawImage, demosaic: int, viewport: Rect):
        g = [float(v) for v in gains.detach().float().cpu().tolist()]
        m = _multiplier(float(stops.detach().cpu()))
        M = [float(v) for v in matrix.detach().float().reshape(-1).cpu().tolist()]
        plan = Program.from_ops(demosaic, False, [(C.OP_GAINS, g), (C.OP_SCALE, (m,)), (C.OP_MATRIX, M)], [])
        out = torch.empty((viewport.height, viewport.width, 3), dtype=torch.float32, device="cuda")
        a = image.metadata.active_area
        ctx.decisions = None
        if demosaic == C.DEMOSAIC_MENON and a.width > 1 and a.height > 1:
            # record the classifier decisions: the adjoint reads them instead of recomputing
            n = ctypes.c_size_t()
            C.check(C.lib().rawops_menon_decisions_size(image._handle, ctypes.byref(n)))
            ctx.decisions = torch.empty(n.value, dtype=torch.uint8, device="cuda")
            C.check(C.lib().rawops_render_record(plan.handle, image._handle, viewport.c(), _ptr(out), viewport.width,
                                                 None, _ptr(ctx.decisions), _stream()))
        else:
            C.check(C.lib().rawops_render(plan.handle, image._handle, viewport.c(), 0, C.QUALITY_FINAL, _ptr(out),
                                          None, viewport.width, None, _stream()))
        ctx.image, ctx.demosaic, ctx.viewport = image, demosaic, viewport
        ctx.g, ctx.m, ctx.M = g, m, M
        return out

    @staticmethod
    def backward(ctx, gout):
        gout = gout.contiguous().float()
        gp = torch.zeros(13, dtype=torch.float64, device="cuda")
        g = (ctypes.c_float * 3)(*ctx.g)
        M = (ctypes.c_float * 9)(*ctx.M)
        if ctx.decisions is not None:
            C.check(C.lib().rawops_render_backward_recorded(ctx.image._handle, ctx.viewport.c(), g, ctx.m, M,
                                                            _ptr(gout), ctx.viewport.width, None, _ptr(gp),
                                                            _ptr(ctx.decisions), ctx.viewport.c(), _stream()))
        else:
            C.check(C.lib().rawops_render_backward(ctx.image._handle, ctx.demosaic, ctx.viewport.c(), g, ctx.m, M,
                                                   _ptr(gout), ctx.viewport.width, None, _ptr(gp), _stream()))
        g_gains = gp[0:3].float()
        g_stops = (gp[3] * ctx.m * _LN2).float()
        g_matrix = gp[4:13].float().reshape(3, 3)
        return g_gains, g_stops, g_matrix, None, None, None


def render_chain(image: RawImage, gains, stops, matrix, demosaic="menon", viewport: Rect = None):
    """Differentiable RAW render: (H, W, 3) FP32 scene-linear output of
    ``viewport`` (default: the active area).  ``gains`` (3), ``stops`` (scalar)
    and ``matrix`` (3x3, camera -> working, float32) are tensors."""
    d = C.DEMOSAIC_MENON if demosaic in ("menon", C.DEMOSAIC_MENON) else C.DEMOSAIC_BILINEAR
    vp = viewport or image.metadata.active_area
    return RenderChainFn.apply(gains, stops, matrix, image, d, vp)


def render_record(image: RawImage, gains, stops, matrix, viewport: Rect = None):
    """Menon forward of the render chain (FP32 out) that also records the
    classifier decisions: (out (h, w, 3), decision record (uint8 bytes of
    little-endian 32-bit words, 32 per 64 x 32 tile of the viewport: with
    X = x - vx, Y = y - vy the non-green site at sensor (x, y) is bit
    (X // 2) % 32 of word ((Y // 32) * ((vw + 63) // 64) + X // 64) * 32 + Y % 32,
    1 = horizontal))."""
    vp = viewport or image.metadata.active_area
    plan = Program.from_ops(C.DEMOSAIC_MENON, False, [(C.OP_GAINS, [float(v) for v in gains]),
                                                      (C.OP_SCALE, (_multiplier(stops),)),
                                                      (C.OP_MATRIX, [float(v) for v in np.asarray(matrix, np.float32).reshape(-1)])], [])
    out = torch.empty((vp.height, vp.width, 3), dtype=torch.float32, device="cuda")
    n = ctypes.c_size_t()
    C.check(C.lib().rawops_menon_decisions_size(image._handle, ctypes.byref(n)))
    dec = torch.zeros(n.value, dtype=torch.uint8, device="cuda")
    C.check(C.lib().rawops_render_record(plan.handle, image._handle, vp.c(), _ptr(out), vp.width, None, _ptr(dec),
                                         _stream()))
    return out, dec


def render_backward(image: RawImage, gout, gains, stops, matrix, demosaic="menon", viewport: Rect = None,
                    want_raw=True, decisions=None, recorded: Rect = None):
    """Explicit adjoint: (dL/d normalised raw plane (ah, aw) or None, dL/d params[13]).
    With ``decisions`` (Menon, from ``render_record`` of the viewport ``recorded``)
    the classifier is read instead of recomputed where the record covers it."""
    d = C.DEMOSAIC_MENON if demosaic in ("menon", C.DEMOSAIC_MENON) else C.DEMOSAIC_BILINEAR
    vp = viewport or image.metadata.active_area
    a = image.metadata.active_area
    graw = torch.zeros((a.height, a.width), dtype=torch.float32, device="cuda") if want_raw else None
    gp = torch.zeros(13, dtype=torch.float64, device="cuda")
    g = (ctypes.c_float * 3)(*[float(v) for v in gains])
    M = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(matrix, np.float32).reshape(-1)])
    if decisions is not None:
        if d != C.DEMOSAIC_MENON:
            raise ValueError("recorded decisions belong to Menon renders")
        C.check(C.lib().rawops_render_backward_recorded(image._handle, vp.c(), g, _multiplier(stops), M,
                                                        _ptr(gout.contiguous()), vp.width, _ptr(graw), _ptr(gp),
                                                        _ptr(decisions), (recorded or vp).c(), _stream()))
    else:
        C.check(C.lib().rawops_render_backward(image._handle, d, vp.c(), g, _multiplier(stops), M,
                                               _ptr(gout.contiguous()), vp.width, _ptr(graw), _ptr(gp), _stream()))
    return graw, gp


# ------------------------------------------------------------- guided filter --

class GuidedFilterFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, eps, radius: int, space: int):
        p = p.contiguous()
        H, W = p.shape[:2]
        e = float(eps.detach().cpu()) if torch.is_tensor(eps) else float(eps)
        src = _Raster(p, space)
        q = torch.empty_like(p)
        C.check(C.lib().rawops_guided_filter(src.h, C.Rect(0, 0, W, H), 0, radius, e, _empty_plan().handle,
                                             _ptr(q), W, None, 0, None, _stream()))
        ctx.save_for_backward(p)
        ctx.e, ctx.radius, ctx.space, ctx.eps_is_tensor = e, radius, space, torch.is_tensor(eps)
        return q

    @staticmethod
    def backward(ctx, gq):
        (p,) = ctx.saved_tensors
        H, W = p.shape[:2]
        gq = gq.contiguous().float()
        gp = torch.empty_like(p)
        geps = torch.zeros(1, dtype=torch.float64, device="cuda")
        if ctx.radius == 0:
            return gq, (torch.zeros((), device="cuda") if ctx.eps_is_tensor else None), None, None
        src = _Raster(p, ctx.space)
        C.check(C.lib().rawops_guided_backward(src.h, ctx.radius, ctx.e, C.Rect(0, 0, W, H), _ptr(gq), W, _ptr(gp),
                                               _ptr(geps), None, 0, _stream()))
        return gp, (geps[0].float() if ctx.eps_is_tensor else None), None, None


def guided_filter(p, eps, radius: int, space: int = 0):
    """Working-Y guided filter of a (H, W, 3) FP32 tensor (radius 0..32)."""
    return GuidedFilterFn.apply(p, eps, radius, space)


# --------------------------------------------------------------- point ops --

class ExposureFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, stops):
        m = _multiplier(float(stops.detach().cpu()))
        y = _apply_program(x, Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [(C.OP_SCALE, (m,))]))
        ctx.save_for_backward(x)
        ctx.m = m
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        gy = gy.contiguous()
        gx = torch.empty_like(x)
        gs = torch.zeros(1, dtype=torch.float64, device="cuda")
        C.check(C.lib().rawops_exposure_backward(_ptr(x), _ptr(gy), ctx.m, _ptr(gx), _ptr(gs), x.numel(), _stream()))
        return gx, gs[0].float()


class SrgbEncodeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        y = torch.empty_like(x)
        C.check(C.lib().rawops_srgb_encode(_ptr(x), _ptr(y), x.numel(), _stream()))
        ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        gx = torch.empty_like(x)
        C.check(C.lib().rawops_srgb_encode_backward(_ptr(x), _ptr(gy.contiguous()), _ptr(gx), x.numel(), _stream()))
        return gx


class DetailBoostFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, base, strength):
        x, base = x.contiguous(), base.contiguous()
        s = float(strength.detach().cpu())
        y = torch.empty_like(x)
        C.check(C.lib().rawops_detail_boost(_ptr(x), _ptr(base), s, _ptr(y), x.numel(), _stream()))
        ctx.save_for_backward(x, base)
        ctx.s = s
        return y

    @staticmethod
    def backward(ctx, gy):
        x, base = ctx.saved_tensors
        gx, gb = torch.empty_like(x), torch.empty_like(x)
        gs = torch.zeros(1, dtype=torch.float64, device="cuda")
        C.check(C.lib().rawops_detail_boost_backward(_ptr(x), _ptr(base), _ptr(gy.contiguous()), ctx.s, _ptr(gx),
                                                     _ptr(gb), _ptr(gs), x.numel(), _stream()))
        return gx, gb, gs[0].float()


class ToneCurveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, knots_y, knots_x, cubic: bool):
        xs = [float(v) for v in knots_x]
        ys = [float(v) for v in knots_y.detach().double().cpu().tolist()]
        n = len(xs)
        ident = [(0.0, 0.0), (1.0, 1.0)]
        flat = list(zip(xs, ys))
        X = (ctypes.c_double * (n + 6))(*([p[0] for p in flat] + [0.0, 1.0] * 3))
        Y = (ctypes.c_double * (n + 6))(*([p[1] for p in flat] + [0.0, 1.0] * 3))
        off = [0, n, n + 2, n + 4, n + 6]
        plan = Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [(C.OP_CURVES, (), (X, Y, off, int(cubic)))])
        y = _apply_program(x, plan)
        ctx.save_for_backward(x)
        ctx.plan, ctx.n = plan, n
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        gx = torch.empty_like(x)
        gk = torch.zeros(ctx.n, dtype=torch.float64, device="cuda")
        C.check(C.lib().rawops_curve_backward(ctx.plan.handle, _ptr(x), _ptr(gy.contiguous()), x.numel(), _ptr(gx),
                                              _ptr(gk), _stream()))
        return gx, gk.float(), None, None


def exposure(x, stops):
    return ExposureFn.apply(x, stops)


def srgb_encode(x):
    return SrgbEncodeFn.apply(x)


def detail_boost(x, base, strength):
    return DetailBoostFn.apply(x, base, strength)


def tone_curve(x, knots_y, knots_x, cubic=True):
    """Master tone curve (ExtendedCurvesNode master, channels identity)."""
    return ToneCurveFn.apply(x, knots_y, knots_x, cubic)


class FullChainCropsFn(torch.autograd.Function):
    """``full_chain`` over several equal-size viewports of one image as one
    autograd node: the render and the render adjoint per crop, the guided
    filter and its adjoint as one launch over the stacked crops (each crop its
    own image), the pointwise stages (exposure, detail boost, sRGB, their
    adjoints) once over the stack; the parameters read to the
    host and the two op plans built once per call, and the parameter gradients
    of all crops accumulated on the device."""

    @staticmethod
    def forward(ctx, gains, stops, matrix, strength, image: RawImage, viewports, demosaic: int, radius: int,
                eps: float):
        host = torch.cat([gains.detach().float().reshape(-1), stops.detach().float().reshape(1),
                          matrix.detach().float().reshape(-1), strength.detach().float().reshape(1)]).cpu().tolist()
        g, st, M, sb = host[0:3], host[3], host[4:13], host[13]
        m = _multiplier(st)
        render = Program.from_ops(demosaic, False, [(C.OP_GAINS, g), (C.OP_SCALE, (1.0,)), (C.OP_MATRIX, M)], [])
        expose = Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [(C.OP_SCALE, (m,))])
        h, w = viewports[0].height, viewports[0].width
        n = len(viewports)
        lin = torch.empty((n, h, w, 3), dtype=torch.float32, device="cuda")
        out = torch.empty_like(lin)
        lin2, base, boosted = torch.empty_like(lin), torch.empty_like(lin), torch.empty_like(lin)
        if any((vp.width, vp.height) != (w, h) for vp in viewports):
            raise ValueError("full_chain_crops needs viewports of one size")
        for i, vp in enumerate(viewports):
            C.check(C.lib().rawops_render(render.handle, image._handle, vp.c(), 0, C.QUALITY_FINAL, _ptr(lin[i]), None,
                                          w, None, _stream()))
        # the exposure (a radius-0 pass: pointwise) over the crops stacked as one raster
        stack = _Raster(lin.view(n * h, w, 3), 0)
        C.check(C.lib().rawops_guided_filter(stack.h, C.Rect(0, 0, w, n * h), 0, 0, 1.0, expose.handle, _ptr(lin2), w,
                                             None, 0, None, _stream()))
        # the guided filter of every crop in one launch (each crop its own image)
        stack2 = _Raster(lin2.view(n * h, w, 3), 0)
        C.check(C.lib().rawops_guided_filter_stack(stack2.h, n, C.Rect(0, 0, w, h), radius, eps,
                                                   _empty_plan().handle, _ptr(base), w, None, _stream()))
        C.check(C.lib().rawops_detail_boost(_ptr(lin2), _ptr(base), sb, _ptr(boosted), lin2.numel(), _stream()))
        C.check(C.lib().rawops_srgb_encode(_ptr(boosted), _ptr(out), boosted.numel(), _stream()))
        ctx.save_for_backward(lin, lin2, base, boosted)
        ctx.image, ctx.viewports, ctx.demosaic, ctx.radius, ctx.eps = image, viewports, demosaic, radius, eps
        ctx.g, ctx.m, ctx.M, ctx.sb = g, m, M, sb
        ctx.plans = (render, expose)
        return out

    @staticmethod
    def backward(ctx, gy):
        lin, lin2, base, boosted = ctx.saved_tensors
        gy = gy.contiguous().float()
        n, h, w = lin.shape[:3]
        gp = torch.zeros(13, dtype=torch.float64, device="cuda")     # gains, (unused), matrix
        gst = torch.zeros(1, dtype=torch.float64, device="cuda")
        gsb = torch.zeros(1, dtype=torch.float64, device="cuda")
        geps = torch.zeros(1, dtype=torch.float64, device="cuda")
        g3 = (ctypes.c_float * 3)(*ctx.g)
        M9 = (ctypes.c_float * 9)(*ctx.M)
        t1, t2, t3 = torch.empty_like(lin), torch.empty_like(lin), torch.empty_like(lin)
        N = lin.numel()
        C.check(C.lib().rawops_srgb_encode_backward(_ptr(boosted), _ptr(gy), _ptr(t1), N, _stream()))
        C.check(C.lib().rawops_detail_boost_backward(_ptr(lin2), _ptr(base), _ptr(t1), ctx.sb, _ptr(t2), _ptr(t3),
                                                     _ptr(gsb), N, _stream()))
        if ctx.radius:
            stack2 = _Raster(lin2.view(n * h, w, 3), 0)
            C.check(C.lib().rawops_guided_backward_stack(stack2.h, n, ctx.radius, ctx.eps, C.Rect(0, 0, w, h),
                                                         _ptr(t3), w, _ptr(t1), _ptr(geps), _stream()))
            t2.add_(t1)
        else:
            t2.add_(t3)
        C.check(C.lib().rawops_exposure_backward(_ptr(lin), _ptr(t2), ctx.m, _ptr(t1), _ptr(gst), N, _stream()))
        for i, vp in enumerate(ctx.viewports):
            C.check(C.lib().rawops_render_backward(ctx.image._handle, ctx.demosaic, vp.c(), g3, 1.0, M9, _ptr(t1[i]), w,
                                                   None, _ptr(gp), _stream()))
        return (gp[0:3].float(), gst[0].float(), gp[4:13].float().reshape(3, 3), gsb[0].float(),
                None, None, None, None, None)


def full_chain_crops(image: RawImage, viewports, gains, stops, matrix, strength, demosaic="menon", radius=8,
                     eps=1e-3):
    """``full_chain`` of each viewport (all of one size), stacked (n, h, w, 3)."""
    d = C.DEMOSAIC_MENON if demosaic in ("menon", C.DEMOSAIC_MENON) else C.DEMOSAIC_BILINEAR
    return FullChainCropsFn.apply(gains, stops, matrix, strength, image, list(viewports), d, radius, float(eps))


def full_chain(image: RawImage, viewport: Rect, gains, stops, matrix, strength, demosaic="menon", radius=8,
               eps=1e-3):
    """RAW -> demosaic -> WB -> exposure -> 3x3 -> guided-filter detail boost
    -> sRGB encode, all differentiable (the fit workload)."""
    lin = render_chain(image, gains, torch.zeros((), device="cuda"), matrix, demosaic, viewport)
    lin = exposure(lin, stops)
    base = guided_filter(lin, eps, radius, 0)
    return srgb_encode(detail_boost(lin, base, strength))
