"""ctypes binding of the C ABI in include/rawops_hip.h (librawops_hip.so).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C librawops_amd/csrc``).  There is no fallback: if the library is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librawops_hip.so")
_LIB = None

OK, INVALID_ARGUMENT, OUT_OF_RANGE, DOMAIN_ERROR, LENGTH_ERROR, OVERFLOW_ERROR, CANCELLED, LOGIC_ERROR, DEVICE_ERROR = range(9)
# RAWOPS_GUIDED_STATIC_SCRATCH_BYTES: the static workspace of rawops_guided_filter_static
GUIDED_STATIC_SCRATCH_BYTES = 64 << 20

OP_GAINS, OP_SCALE, OP_MATRIX, OP_TONE, OP_CLIP, OP_SRGB_ENCODE, OP_CURVES = range(1, 8)
DEMOSAIC_BILINEAR, DEMOSAIC_MENON = 0, 1
QUALITY_PREVIEW, QUALITY_FINAL = 0, 1


class RawOpsError(RuntimeError):
    """Base of the errors raised for a non-OK rawops_status."""


class InvalidArgument(RawOpsError, ValueError):
    """std::invalid_argument."""


class OutOfRange(RawOpsError, IndexError):
    """std::out_of_range (viewport outside the source)."""


class DomainError(RawOpsError, ArithmeticError):
    """std::domain_error (non-finite value at an output boundary)."""


class LengthError(RawOpsError, MemoryError):
    """std::length_error."""


class FilterOverflow(RawOpsError, OverflowError):
    """std::overflow_error (guided filter)."""


class RenderCancelled(RawOpsError):
    """rawengine::RenderCancelled."""


class LogicError(RawOpsError):
    """std::logic_error."""


class DeviceError(RawOpsError):
    """HIP runtime failure."""


_ERRORS = {INVALID_ARGUMENT: InvalidArgument, OUT_OF_RANGE: OutOfRange, DOMAIN_ERROR: DomainError,
           LENGTH_ERROR: LengthError, OVERFLOW_ERROR: FilterOverflow, CANCELLED: RenderCancelled,
           LOGIC_ERROR: LogicError, DEVICE_ERROR: DeviceError}


class Rect(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]

    def tuple(self):
        return (self.x, self.y, self.width, self.height)


class RawMetadataC(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("row_stride_samples", ctypes.c_uint32), ("pattern", ctypes.c_uint32),
                ("cfa_phase_x", ctypes.c_uint32), ("cfa_phase_y", ctypes.c_uint32),
                ("active_area", Rect),
                ("black_levels", ctypes.c_uint16 * 4), ("white_levels", ctypes.c_uint16 * 4)]


class OpDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("v", ctypes.c_float * 9),
                ("knots_x", ctypes.POINTER(ctypes.c_double)),
                ("knots_y", ctypes.POINTER(ctypes.c_double)),
                ("knot_offsets", ctypes.c_uint32 * 5), ("cubic", ctypes.c_uint32)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("demosaic", ctypes.c_uint32), ("mip_anchor", ctypes.c_uint32),
                ("n_native", ctypes.c_uint32), ("n_reduced", ctypes.c_uint32),
                ("native_ops", ctypes.POINTER(OpDesc)), ("reduced_ops", ctypes.POINTER(OpDesc))]


class BatchRequestC(ctypes.Structure):
    _fields_ = [("plan", ctypes.c_void_p), ("source", ctypes.c_void_p), ("viewport", Rect),
                ("mip", ctypes.c_uint32), ("quality", ctypes.c_uint32), ("linear_ops", ctypes.c_uint32),
                ("dst_rgb", ctypes.c_void_p), ("dst_u8", ctypes.c_void_p), ("dst_stride_pixels", ctypes.c_uint32)]


TILE_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, Rect, ctypes.POINTER(ctypes.c_float))


def lib():
    """Load librawops_hip.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or make -C librawops_amd/csrc (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        P, S, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32
        PP = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "rawops_last_error": ([], ctypes.c_char_p),
            "rawops_version": ([], ctypes.c_char_p),
            "rawops_camera_to_working_matrix": ([P, u32, P], S),
            "rawops_working_to_srgb_matrix": ([u32, P], S),
            "rawops_working_to_working_matrix": ([u32, u32, P], S),
            "rawops_exposure_multiplier": ([ctypes.c_float, P], S),
            "rawops_xyz_d50_to_working_matrix": ([u32, P], S),
            "rawops_source_create_raw": ([P, P, ctypes.c_size_t, PP], S),
            "rawops_source_wrap_raw_device": ([P, P, ctypes.c_size_t, PP], S),
            "rawops_source_create_raster": ([u32, u32, u32, u32, P, PP], S),
            "rawops_source_wrap_raster_device": ([u32, u32, u32, u32, P, PP], S),
            "rawops_source_raw_metadata": ([P, P], S),
            "rawops_source_destroy": ([P], None),
            "rawops_plan_create": ([P, PP], S),
            "rawops_plan_destroy": ([P], None),
            "rawops_request_bounds": ([Rect, u32, u32, P], S),
            "rawops_render": ([P, P, Rect, u32, u32, P, P, u32, P, P], S),
            "rawops_check_flags": ([P, P], S),
            "rawops_render_tiles_host": ([P, P, Rect, u32, u32, u32, TILE_CALLBACK, P, P], S),
            "rawops_batch_create": ([P, ctypes.c_size_t, u32, P, PP], S),
            "rawops_batch_run": ([P, P], S),
            "rawops_batch_latencies": ([P, P], S),
            "rawops_batch_check": ([P, P], S),
            "rawops_batch_shape": ([P, P, P, P], S),
            "rawops_batch_shape_count": ([], u32),
            "rawops_batch_destroy": ([P], None),
            "rawops_render_backward": ([P, u32, Rect, P, ctypes.c_float, P, P, u32, P, P, P], S),
            "rawops_menon_decisions_size": ([P, P], S),
            "rawops_render_record": ([P, P, Rect, P, u32, P, P, P], S),
            "rawops_render_backward_recorded": ([P, Rect, P, ctypes.c_float, P, P, u32, P, P, P, Rect, P], S),
            "rawops_guided_backward_workspace_size": ([u32, u32, P], S),
            "rawops_guided_backward": ([P, u32, ctypes.c_double, Rect, P, u32, P, P, P, ctypes.c_size_t, P], S),
            "rawops_srgb_encode": ([P, P, ctypes.c_size_t, P], S),
            "rawops_srgb_encode_backward": ([P, P, P, ctypes.c_size_t, P], S),
            "rawops_exposure_backward": ([P, P, ctypes.c_float, P, P, ctypes.c_size_t, P], S),
            "rawops_detail_boost": ([P, P, ctypes.c_float, P, ctypes.c_size_t, P], S),
            "rawops_detail_boost_backward": ([P, P, P, ctypes.c_float, P, P, P, ctypes.c_size_t, P], S),
            "rawops_curve_backward": ([P, P, P, ctypes.c_size_t, P, P, P], S),
            "rawops_exposure_curve_backward": ([P, P, ctypes.c_float, P, ctypes.c_size_t, P, P, P, P], S),
            "rawops_guided_workspace_size": ([u32, u32, Rect, u32, u32, P], S),
            "rawops_guided_filter": ([P, Rect, u32, u32, ctypes.c_double, P, P, u32, P,
                                      ctypes.c_size_t, P, P], S),
            "rawops_set_menon_path": ([u32, u32], S),
            "rawops_guided_filter_static": ([P, Rect, u32, u32, ctypes.c_double, P, P, u32, P, P], S),
            "rawops_histogram_rgb": ([P, u32, u32, u32, u32, ctypes.c_double, ctypes.c_double, P, P, P, P], S),
            "rawops_guided_filter_stack": ([P, u32, Rect, u32, ctypes.c_double, P, P, u32, P, P], S),
            "rawops_guided_backward_stack": ([P, u32, u32, ctypes.c_double, Rect, P, u32, P, P, P], S),
            "rawops_local_statistics": ([P, Rect, u32, Rect, Rect, u32, P, P, P, P], S),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(status):
    """Raise the exception matching a rawops_status."""
    if status != OK:
        msg = lib().rawops_last_error().decode()
        raise _ERRORS.get(status, RawOpsError)(msg)


def exported_symbols():
    """Names declared in include/rawops_hip.h that the library must export."""
    import re
    header = os.path.join(os.path.dirname(_HERE), "include", "rawops_hip.h")
    text = open(header).read()
    return sorted(set(re.findall(r"\b(rawops_[a-z0-9_]+)\s*\(", text)))


def local_statistics(tile, needed, output, image, radius, stream=None):
    """rawops_local_statistics over a device float32 tensor (h, w, 3) covering
    ``needed``: (mean, variance, finite_count) device tensors (h, w, 3)."""
    import torch
    assert tile.is_cuda and tile.dtype == torch.float32 and tile.is_contiguous()
    assert tuple(tile.shape) == (needed[3], needed[2], 3)
    shape = (output[3], output[2], 3)
    mean = torch.empty(shape, dtype=torch.float64, device=tile.device)
    var = torch.empty(shape, dtype=torch.float64, device=tile.device)
    cnt = torch.empty(shape, dtype=torch.int32, device=tile.device)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().rawops_local_statistics(tile.data_ptr(), Rect(*needed), needed[2], Rect(*output), Rect(*image),
                                        radius, mean.data_ptr(), var.data_ptr(), cnt.data_ptr(), s))
    torch.cuda.synchronize()
    return mean, var, cnt
