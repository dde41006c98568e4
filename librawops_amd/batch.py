"""Batched heterogeneous render: many RAW render requests in one
hipGraph-captured launch, one workgroup per (request, tile) item, the tile
shape autotuned on the batch (see include/rawops_hip.h, rawops_batch_*).

Each request is a node chain (compiled like Renderer's), a viewport at a
level, and device output tensors: an FP32 RGB tensor taken after the first
``linear_ops`` reduced ops of the chain, and/or a uint8 tensor of the final
values.  For an SrgbPreview graph with camera calibration, ``linear_ops=0``
yields FP32 scene-linear output plus the uint8 sRGB preview from one pass.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import torch

from . import _capi as C
from .graph import Node, RenderLevel, Rect, compile_chain


@dataclass
class BatchRequest:
    node: Node
    viewport: Rect
    level: RenderLevel = RenderLevel()
    linear_ops: int = 0
    out_rgb: Optional[torch.Tensor] = None
    out_u8: Optional[torch.Tensor] = None


class Batch:
    """A captured batch; ``run()`` replays it on the current stream."""

    def __init__(self, requests, shape: Optional[int] = None, stream=None):
        self.requests = list(requests)
        self._compiled = [compile_chain(r.node) for r in self.requests]
        arr = (C.BatchRequestC * len(self.requests))()
        for i, (r, c) in enumerate(zip(self.requests, self._compiled)):
            if c.kind != "raw":
                raise C.InvalidArgument("batched requests must render RAW sources")
            arr[i].plan = c.program.handle
            arr[i].source = c.source.image._handle
            arr[i].viewport = r.viewport.c()
            arr[i].mip = r.level.mip
            arr[i].quality = int(r.level.quality)
            arr[i].linear_ops = r.linear_ops
            arr[i].dst_rgb = r.out_rgb.data_ptr() if r.out_rgb is not None else None
            arr[i].dst_u8 = r.out_u8.data_ptr() if r.out_u8 is not None else None
            arr[i].dst_stride_pixels = 0
        self._arr = arr
        s = stream or torch.cuda.current_stream()
        h = ctypes.c_void_p()
        C.check(C.lib().rawops_batch_create(arr, len(self.requests), 0xFFFFFFFF if shape is None else shape,
                                            ctypes.c_void_p(s.cuda_stream), ctypes.byref(h)))
        self._h = h
        self._free = C.lib().rawops_batch_destroy   # bound now: usable during interpreter exit

    def run(self, stream=None):
        s = stream or torch.cuda.current_stream()
        C.check(C.lib().rawops_batch_run(self._h, ctypes.c_void_p(s.cuda_stream)))

    def check(self, stream=None):
        s = stream or torch.cuda.current_stream()
        C.check(C.lib().rawops_batch_check(self._h, ctypes.c_void_p(s.cuda_stream)))

    def latencies_ms(self):
        out = (ctypes.c_double * len(self.requests))()
        C.check(C.lib().rawops_batch_latencies(self._h, out))
        return list(out)

    def shape(self):
        tw, th, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        C.check(C.lib().rawops_batch_shape(self._h, ctypes.byref(tw), ctypes.byref(th), ctypes.byref(ms)))
        return tw.value, th.value, ms.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._free(h)
            self._h = None
