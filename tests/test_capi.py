"""CPU-side checks of the C ABI boundary: the library loads, exports every
entry point include/rawops_hip.h declares, and its host-only colour math
matches the (pinned) oracle bit for bit.  No device call is made here."""
import ctypes
import subprocess

import numpy as np
import pytest

from librawops_amd import _capi as C
from oracle import oracle as O

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = C.exported_symbols()
    assert "rawops_render" in names and "rawops_guided_filter" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert f" {n}\n" in nm, n


def _f9(fn, *args):
    out = (ctypes.c_float * 9)()
    C.check(fn(*args, out))
    return np.array(list(out), np.float32)


@pytest.mark.parametrize("space", [0, 1])
def test_host_matrices_bit_exact(space):
    cam = (ctypes.c_double * 9)(*CAM)
    got = _f9(C.lib().rawops_camera_to_working_matrix, cam, space)
    assert got.tobytes() == O.camera_to_working_f32(CAM, space).tobytes()
    assert _f9(C.lib().rawops_working_to_srgb_matrix, space).tobytes() == O.working_to_linear_srgb_f32(space).tobytes()
    assert (_f9(C.lib().rawops_working_to_working_matrix, space, 1 - space).tobytes()
            == O.working_to_working_f32(space, 1 - space).tobytes())


def test_host_validation_statuses():
    cam = (ctypes.c_double * 9)(1, 0, 0, 1, 0, 0, 0, 0, 1)   # singular
    out = (ctypes.c_float * 9)()
    assert C.lib().rawops_camera_to_working_matrix(cam, 0, out) == C.INVALID_ARGUMENT
    assert b"singular" in C.lib().rawops_last_error()
    m = ctypes.c_float()
    assert C.lib().rawops_exposure_multiplier(ctypes.c_float(33.0), ctypes.byref(m)) == C.INVALID_ARGUMENT
    for stops in (-2.5, 0.0, 0.7, 3.0):
        C.check(C.lib().rawops_exposure_multiplier(ctypes.c_float(stops), ctypes.byref(m)))
        assert np.float32(m.value) == O.exp2f(stops)
    r = C.Rect()
    C.check(C.lib().rawops_request_bounds(C.Rect(16, 16, 6016, 3991), 2, 0, ctypes.byref(r)))
    assert r.tuple() == (0, 0, 1504, 998)
    assert C.lib().rawops_request_bounds(C.Rect(0, 0, 4, 4), 1, 1, ctypes.byref(r)) == C.INVALID_ARGUMENT
    assert C.lib().rawops_request_bounds(C.Rect(0, 0, 4, 4), 3, 0, ctypes.byref(r)) == C.INVALID_ARGUMENT


def test_metadata_validation_without_device():
    """RawImage ctor rules are checked before any device allocation."""
    m = C.RawMetadataC()
    m.width, m.height = 4, 4
    for i in range(4):
        m.white_levels[i] = 100
    h = ctypes.c_void_p()
    buf = (ctypes.c_uint16 * 15)()
    assert C.lib().rawops_source_create_raw(ctypes.byref(m), buf, 15, ctypes.byref(h)) == C.INVALID_ARGUMENT
    m.pattern = 4
    assert C.lib().rawops_source_create_raw(ctypes.byref(m), buf, 16, ctypes.byref(h)) == C.INVALID_ARGUMENT
    m.pattern = 0
    m.active_area = C.Rect(2, 2, 3, 3)
    assert C.lib().rawops_source_create_raw(ctypes.byref(m), buf, 16, ctypes.byref(h)) == C.INVALID_ARGUMENT
    m.active_area = C.Rect()
    m.black_levels[2] = 100
    assert C.lib().rawops_source_create_raw(ctypes.byref(m), buf, 16, ctypes.byref(h)) == C.INVALID_ARGUMENT
