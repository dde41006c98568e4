"""rawops_guided_filter_static, the entry point for callers that cannot
allocate (its per-centre coefficient workspace is a static device array, so
the call itself allocates nothing), through the C ABI against
rawops_guided_filter with a caller workspace: bit-identical outputs on two
calls in a row, the second on a smaller input (the scratch keeps the first
call's coefficients, which the second must fully overwrite where it reads),
and a workspace beyond the static size rejected with RAWOPS_LENGTH_ERROR
before any launch."""
import ctypes

import numpy as np
import pytest
import torch

import librawops_amd as L
from librawops_amd import _capi as C
from librawops_amd import synth
from librawops_amd.graph import Program

pytestmark = pytest.mark.gpu


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def post_plan(stops, knots):
    """Exposure then a 16-knot cubic master curve (identity channel curves)."""
    m = ctypes.c_float()
    C.check(C.lib().rawops_exposure_multiplier(ctypes.c_float(stops), ctypes.byref(m)))
    kx = [k[0] for k in knots]
    ky = [k[1] for k in knots]
    X = (ctypes.c_double * 22)(*(kx + [0.0, 1.0] * 3))
    Y = (ctypes.c_double * 22)(*(ky + [0.0, 1.0] * 3))
    return Program.from_ops(C.DEMOSAIC_BILINEAR, False, [],
                            [(C.OP_SCALE, (m.value,)), (C.OP_CURVES, (), (X, Y, [0, 16, 18, 20, 22], 1))])


def render_pair(width, height, vp, mip, radius, eps, seed):
    raster = torch.from_numpy(synth.lognormal_raster(width, height, seed=seed)).cuda()
    img = L.RasterImage(L.RasterMetadata(width, height, 0, L.WorkingSpace.LinearRec2020D65), raster)
    plan = post_plan(0.7, synth.monotone_knots(16, seed=seed))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rect = L.Rect(*vp)
    need = ctypes.c_size_t()
    C.check(C.lib().rawops_guided_workspace_size(width, height, rect.c(), mip, radius, ctypes.byref(need)))
    assert 0 < need.value <= C.GUIDED_STATIC_SCRATCH_BYTES
    ws = torch.empty(need.value, dtype=torch.uint8, device="cuda")
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    want = torch.full((vp[3], vp[2], 3), float("nan"), device="cuda")
    got = torch.full((vp[3], vp[2], 3), float("nan"), device="cuda")
    C.check(C.lib().rawops_guided_filter(img._handle, rect.c(), mip, radius, eps, plan.handle, P(want), vp[2],
                                         P(ws), ws.numel(), P(flags), stream))
    print(f"no-alloc: static call {width}x{height} vp {vp} mip {mip} r {radius}", flush=True)
    C.check(C.lib().rawops_guided_filter_static(img._handle, rect.c(), mip, radius, eps, plan.handle, P(got), vp[2],
                                                ctypes.c_void_p(flags.data_ptr() + 4), stream))
    torch.cuda.synchronize()
    assert int(flags[0]) == 0 and int(flags[1]) == 0
    return got, want


def test_static_scratch_matches_allocating_entry_point():
    # the larger call first, then a smaller one over the scratch it left
    for width, height, vp, mip, radius, eps, seed in [(1500, 1100, (0, 0, 1500, 1100), 0, 8, 1e-3, 31),
                                                      (400, 300, (17, 9, 171, 121), 1, 4, 1e-2, 32)]:
        got, want = render_pair(width, height, vp, mip, radius, eps, seed)
        assert not torch.isnan(want).any()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (width, height, vp, mip, radius)


def test_static_scratch_rejects_oversize_workspace():
    width, height, radius = 2000, 1500, 8
    raster = torch.zeros((height, width, 3), dtype=torch.float32, device="cuda")
    img = L.RasterImage(L.RasterMetadata(width, height, 0, L.WorkingSpace.LinearRec2020D65), raster)
    plan = Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], [])
    rect = L.Rect(0, 0, width, height)
    need = ctypes.c_size_t()
    C.check(C.lib().rawops_guided_workspace_size(width, height, rect.c(), 0, radius, ctypes.byref(need)))
    assert need.value > C.GUIDED_STATIC_SCRATCH_BYTES
    out = torch.full((height, width, 3), float("nan"), device="cuda")
    st = C.lib().rawops_guided_filter_static(img._handle, rect.c(), 0, radius, 1e-3, plan.handle, P(out), width, None,
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == C.LENGTH_ERROR
    assert b"static scratch" in C.lib().rawops_last_error()
    torch.cuda.synchronize()
    assert torch.isnan(out).all()
