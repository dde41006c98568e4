"""histogram_rgb (ImageAnalysis.hpp:7-40, ImageAnalysis.cpp:10-80): a
per-channel histogram of exactly the requested graph output, computed on the
device by one histogram pass (rawops_histogram_rgb) over the device render of
the viewport; only the counters come back to the host."""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _capi as C
from .graph import (ImageDescriptor, ImageGraph, InvalidArgument, RenderCancelled, RenderRequest, Renderer,
                    _validate_rect, request_bounds)


@dataclass
class HistogramOptions:
    """HistogramOptions (ImageAnalysis.hpp:7-11): bins 1..65536, equally spaced in [lower, upper]."""
    bins: int = 256
    lower: float = 0.0
    upper: float = 1.0


@dataclass
class ChannelHistogram:
    """ChannelHistogram (ImageAnalysis.hpp:13-17)."""
    counts: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    underflow: int = 0
    overflow: int = 0
    nonfinite: int = 0
    minimum: Optional[float] = None     # all finite values, including outliers
    maximum: Optional[float] = None


@dataclass
class RgbHistogram:
    """RgbHistogram (ImageAnalysis.hpp:19-25)."""
    request: RenderRequest
    options: HistogramOptions
    descriptor: ImageDescriptor
    pixel_count: int = 0
    channels: List[ChannelHistogram] = field(default_factory=list)


def _key_float(key: int) -> float:
    """Inverse of the device order key: bits ^ 0x80000000 for v >= +0, ~bits for v <= -0."""
    bits = (key & 0x7FFFFFFF) if key & 0x80000000 else (~key & 0xFFFFFFFF)
    return float(np.array([bits], np.uint32).view(np.float32)[0])


def histogram_rgb(target, *args, options: HistogramOptions = None, cancellation=None) -> RgbHistogram:
    """histogram_rgb(graph, request[, options]) or (node, source_bounds, request[, options]).

    Bins are half-open except the last, which includes ``upper``; values
    outside [lower, upper] are tallied as underflow / overflow, not clipped;
    NaN and infinities count as nonfinite and do not enter the extrema
    (ImageAnalysis.cpp:52-63).  Raises as the reference does: InvalidArgument
    for bad options (before any rendering), an empty or unaddressable
    viewport, a zero tile size or an unsupported level; OutOfRange for a
    viewport outside the source; RenderCancelled when cancelled."""
    if isinstance(target, ImageGraph):
        node, bounds, rest = target.output(), target.source_bounds(), list(args)
    else:
        node, bounds, rest = target, args[0], list(args[1:])
    request = rest[0]
    if options is None:
        options = rest[1] if len(rest) > 1 else HistogramOptions()
    bins, lower, upper = int(options.bins), float(options.lower), float(options.upper)
    span = upper - lower
    if not (1 <= bins <= 65536) or not math.isfinite(lower) or not math.isfinite(upper) or not (span > 0.0) \
            or not math.isfinite(span):
        raise InvalidArgument("histogram requires 1..65536 bins and a finite positive range")
    if cancellation is not None and cancellation.is_cancelled():
        raise RenderCancelled("render cancelled")
    v = request.viewport
    if not v.width or not v.height or v.x + v.width > (1 << 32) or v.y + v.height > (1 << 32):
        raise InvalidArgument("histogram requires a nonempty addressable viewport")
    _validate_rect(request_bounds(bounds, request.level), v)
    if not node.supports_level(request.level):
        raise InvalidArgument("node does not support this render level")
    if not request.tile_size:
        raise InvalidArgument("callback and tile size are required")
    import torch
    counts = torch.zeros(3 * bins, dtype=torch.int64, device="cuda")
    tallies = torch.zeros(9, dtype=torch.int64, device="cuda")
    extrema = torch.tensor([-1, 0] * 3, dtype=torch.int32, device="cuda")    # keys 0xFFFFFFFF, 0
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # Bands of whole rows (at least a tile high, up to ~8 M pixels), each
    # rendered on the device and added into the same counters: the histogram
    # does not depend on how the viewport is partitioned, and the device
    # staging stays one band, not the whole viewport.
    band_h = min(v.height, max(int(request.tile_size), (8 << 20) // v.width + 1))
    y = v.y
    while y < v.y + v.height:
        if cancellation is not None and cancellation.is_cancelled():
            raise RenderCancelled("render cancelled")
        band = type(v)(v.x, y, v.width, min(band_h, v.y + v.height - y))
        image = Renderer()._render_viewport_device(node, band, request.level)
        C.check(C.lib().rawops_histogram_rgb(ctypes.c_void_p(image.data_ptr()), band.width, band.height, band.width,
                                             bins, lower, upper, ctypes.c_void_p(counts.data_ptr()),
                                             ctypes.c_void_p(tallies.data_ptr()),
                                             ctypes.c_void_p(extrema.data_ptr()), stream))
        y += band.height
    if cancellation is not None and cancellation.is_cancelled():
        raise RenderCancelled("render cancelled")
    c = counts.cpu().numpy().view(np.uint64).reshape(3, bins)
    t = tallies.cpu().numpy().view(np.uint64).reshape(3, 3)
    e = extrema.cpu().numpy().view(np.uint32).reshape(3, 2)
    channels = []
    for k in range(3):
        kmin, kmax = int(e[k, 0]), int(e[k, 1])
        finite = kmin != 0xFFFFFFFF
        channels.append(ChannelHistogram(c[k].copy(), int(t[k, 0]), int(t[k, 1]), int(t[k, 2]),
                                         _key_float(kmin) if finite else None, _key_float(kmax) if finite else None))
    return RgbHistogram(request, options, node.output_descriptor(), int(v.width) * int(v.height), channels)
