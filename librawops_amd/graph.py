"""Python mirror of the reference's render-path API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference's C++ API
for the tiled render path (RawEngine.hpp:26-710; SpatialOps.hpp:42-69;
ToneOps.hpp:7-50): ``RawMetadata``/``RawImage``, the node classes,
``ImageGraph``, ``RenderRequest`` and ``Renderer.render_tiles/render_image``.

Nodes are immutable descriptions.  Rendering compiles the chain that ends at
the requested node into a fused device plan (demosaic + per-pixel op program
+ mip reduction, or guided filter + op program) and runs it on the GPU; the
source samples live in HBM for the life of the ``RawImage``/``RasterImage``.
There is no CPU execution path: without the HIP library every call raises.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Callable, Optional

import numpy as np

from . import _capi as C
from ._capi import (DomainError, FilterOverflow, InvalidArgument, OutOfRange,  # noqa: F401
                    RawOpsError, RenderCancelled)


class BayerPattern(IntEnum):
    RGGB = 0
    BGGR = 1
    GRBG = 2
    GBRG = 3


class WorkingSpace(IntEnum):
    LinearProPhotoD50 = 0
    LinearRec2020D65 = 1


class RenderQuality(IntEnum):
    Preview = 0
    Final = 1


class OutputMode(IntEnum):
    LegacyBounded = 0
    SrgbPreview = 1
    IccDisplay = 2


class CurveInterpolation(IntEnum):
    Linear = 0
    ShapePreservingCubic = 1


@dataclass(frozen=True)
class Rect:
    x: int = 0
    y: int = 0
    width: int = 0
    height: int = 0

    def c(self):
        return C.Rect(self.x, self.y, self.width, self.height)


@dataclass(frozen=True)
class RenderLevel:
    mip: int = 0
    quality: RenderQuality = RenderQuality.Final


@dataclass(frozen=True)
class RenderRequest:
    viewport: Rect
    tile_size: int = 256
    level: RenderLevel = RenderLevel()


@dataclass(frozen=True)
class ImageDescriptor:
    """ImageDescriptor (RawEngine.hpp:154-204), as a comparable value."""
    domain: str = "camera_linear"
    primaries: str = "camera"
    white_point: str = "unspecified"
    transfer: str = "linear"

    @staticmethod
    def camera_linear():
        return ImageDescriptor()

    @staticmethod
    def scene_linear(space):
        if WorkingSpace(space) == WorkingSpace.LinearProPhotoD50:
            return ImageDescriptor("scene_linear", "prophoto", "d50", "linear")
        return ImageDescriptor("scene_linear", "rec2020", "d65", "linear")

    @staticmethod
    def linear_srgb():
        return ImageDescriptor("scene_linear", "srgb", "d65", "linear")

    @staticmethod
    def tone_mapped(primaries="camera", white_point="unspecified"):
        return ImageDescriptor("tone_mapped", primaries, white_point, "custom_tone")

    @staticmethod
    def display_linear_srgb():
        return ImageDescriptor("display_linear", "srgb", "d65", "linear")

    @staticmethod
    def srgb_output():
        return ImageDescriptor("display_encoded", "srgb", "d65", "srgb")


SCENE = (ImageDescriptor.scene_linear(0), ImageDescriptor.scene_linear(1))


@dataclass
class Tile:
    bounds: Rect
    rgb: np.ndarray                  # float32, (height, width, 3)
    descriptor: ImageDescriptor = ImageDescriptor()


@dataclass
class RawMetadata:
    width: int = 0
    height: int = 0
    row_stride_samples: int = 0
    pattern: BayerPattern = BayerPattern.RGGB
    cfa_phase_x: int = 0
    cfa_phase_y: int = 0
    active_area: Rect = Rect()
    black_levels: tuple = (0, 0, 0, 0)
    white_levels: tuple = (65535, 65535, 65535, 65535)

    def c(self):
        m = C.RawMetadataC()
        m.width, m.height, m.row_stride_samples = self.width, self.height, self.row_stride_samples
        m.pattern, m.cfa_phase_x, m.cfa_phase_y = int(self.pattern), self.cfa_phase_x, self.cfa_phase_y
        m.active_area = self.active_area.c()
        for i in range(4):
            m.black_levels[i] = int(self.black_levels[i])
            m.white_levels[i] = int(self.white_levels[i])
        return m


@dataclass
class RasterMetadata:
    width: int = 0
    height: int = 0
    row_stride_pixels: int = 0
    working_space: WorkingSpace = WorkingSpace.LinearProPhotoD50


@dataclass(frozen=True)
class RawDemosaicIdentity:
    algorithm: str = "rawengine.bilinear"
    processing_version: int = 1


def validate_raw_demosaic(identity: RawDemosaicIdentity):
    """RawEngine.cpp:276-280."""
    if identity.algorithm not in ("rawengine.bilinear", "rawengine.menon_base") or identity.processing_version != 1:
        raise InvalidArgument("unsupported RAW demosaic algorithm or processing version")


@dataclass
class CameraColorTransform:
    camera_to_xyz_d50: tuple = (0.0,) * 9
    target: WorkingSpace = WorkingSpace.LinearProPhotoD50


@dataclass
class GraphRecipe:
    red_gain: float = 1.0
    green_gain: float = 1.0
    blue_gain: float = 1.0
    exposure_stops: float = 0.0
    tone_shoulder: float = 0.25
    tone_gamma: float = 1.0
    camera_color: Optional[CameraColorTransform] = None
    output_mode: OutputMode = OutputMode.LegacyBounded


@dataclass
class CurvePoint:
    x: float
    y: float


@dataclass
class ExtendedCurvesSettings:
    master: list = field(default_factory=lambda: [(0.0, 0.0), (1.0, 1.0)])
    channels: list = field(default_factory=lambda: [[(0.0, 0.0), (1.0, 1.0)] for _ in range(3)])
    interpolation: CurveInterpolation = CurveInterpolation.Linear


@dataclass
class WorkingYGuidedFilterSettings:
    radius: int = 3
    epsilon: float = 2.0 ** -12


class CancellationToken:
    def __init__(self):
        self._flag = ctypes.c_int(0)

    def cancel(self):
        self._flag.value = 1

    def is_cancelled(self):
        return bool(self._flag.value)


# ------------------------------------------------------------------ sources --

def _torch():
    import torch
    return torch


class RawImage:
    """Decoded uint16 Bayer samples, device-resident (RawEngine.hpp:70-92).

    ``samples`` is a host array (uploaded once) or a CUDA tensor (wrapped
    without a copy; the caller keeps it alive)."""

    def __init__(self, metadata: RawMetadata, samples):
        self._keep = None
        h = ctypes.c_void_p()
        m = metadata.c()
        if hasattr(samples, "is_cuda") and samples.is_cuda:
            t = _torch()
            if samples.dtype not in (t.uint16, t.int16):
                raise InvalidArgument("device samples must be 16-bit")
            samples = samples.contiguous()
            self._keep = samples
            C.check(C.lib().rawops_source_wrap_raw_device(ctypes.byref(m), ctypes.c_void_p(samples.data_ptr()),
                                                          samples.numel(), ctypes.byref(h)))
        else:
            s = np.ascontiguousarray(samples, dtype=np.uint16).reshape(-1)
            C.check(C.lib().rawops_source_create_raw(ctypes.byref(m), s.ctypes.data_as(ctypes.c_void_p),
                                                     s.size, ctypes.byref(h)))
        self._handle = h
        self._free = C.lib().rawops_source_destroy   # bound now: usable during interpreter exit
        out = C.RawMetadataC()
        C.check(C.lib().rawops_source_raw_metadata(h, ctypes.byref(out)))
        a = out.active_area
        self.metadata = RawMetadata(out.width, out.height, out.row_stride_samples, BayerPattern(out.pattern),
                                    out.cfa_phase_x, out.cfa_phase_y, Rect(a.x, a.y, a.width, a.height),
                                    tuple(out.black_levels), tuple(out.white_levels))

    @property
    def width(self):
        return self.metadata.width

    @property
    def height(self):
        return self.metadata.height

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._free(h)
            self._handle = None


class RasterImage:
    """Scene-linear interleaved float32 RGB (RawEngine.hpp:94-119)."""

    def __init__(self, metadata: RasterMetadata, pixels):
        self._keep = None
        h = ctypes.c_void_p()
        M = metadata
        if hasattr(pixels, "is_cuda") and pixels.is_cuda:
            p = pixels.contiguous().float()
            self._keep = p
            C.check(C.lib().rawops_source_wrap_raster_device(M.width, M.height, M.row_stride_pixels,
                                                             int(M.working_space), ctypes.c_void_p(p.data_ptr()),
                                                             ctypes.byref(h)))
        else:
            p = np.ascontiguousarray(pixels, dtype=np.float32).reshape(-1)
            stride = M.row_stride_pixels or M.width
            if p.size != stride * M.height * 3:
                raise InvalidArgument("raster float count does not match stride and height")
            C.check(C.lib().rawops_source_create_raster(M.width, M.height, M.row_stride_pixels,
                                                        int(M.working_space), p.ctypes.data_as(ctypes.c_void_p),
                                                        ctypes.byref(h)))
        self._handle = h
        self._free = C.lib().rawops_source_destroy   # bound now: usable during interpreter exit
        self.metadata = RasterMetadata(M.width, M.height, M.row_stride_pixels or M.width, WorkingSpace(M.working_space))

    @property
    def width(self):
        return self.metadata.width

    @property
    def height(self):
        return self.metadata.height

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._free(h)
            self._handle = None


# -------------------------------------------------------------------- nodes --

def _final(level):
    return level.mip == 0 and level.quality == RenderQuality.Final


def _reduced(level):
    return 1 <= level.mip <= 2 and level.quality == RenderQuality.Preview


def request_bounds(source_bounds: Rect, level: RenderLevel) -> Rect:
    """RawEngine.cpp:31-43."""
    if level.mip == 0 and level.quality in (RenderQuality.Final, RenderQuality.Preview):
        return source_bounds
    if _reduced(level):
        s = 1 << level.mip
        return Rect(0, 0, -(-source_bounds.width // s), -(-source_bounds.height // s))
    raise InvalidArgument("render level is not implemented for these source bounds")


class Node:
    """The reference's Node interface (RawEngine.hpp:232-269)."""
    input: Optional["Node"] = None

    def input_node(self):
        return self.input

    def output_descriptor(self) -> ImageDescriptor:
        raise NotImplementedError

    def supports_level(self, level: RenderLevel) -> bool:
        return _final(level)

    def source_bounds(self):
        return None

    def render(self, bounds: Rect) -> Tile:
        return self.render_level(bounds, RenderLevel())

    def render_level(self, bounds: Rect, level: RenderLevel) -> Tile:
        if not self.supports_level(level):
            raise InvalidArgument("node does not support this render level")
        src = _source_of(self)
        return Renderer().render_image(self, src.source_bounds(), RenderRequest(bounds, max(1, bounds.width, bounds.height), level))

    # Op program contribution: list of (kind, values, extra) for point nodes.
    def _ops(self):
        return []


def _source_of(node):
    while node.input_node() is not None:
        node = node.input_node()
    return node


class RawUnpackNode(Node):
    def __init__(self, image: RawImage, demosaic: RawDemosaicIdentity = RawDemosaicIdentity()):
        validate_raw_demosaic(demosaic)
        self.image = image
        self.demosaic = demosaic

    def output_descriptor(self):
        return ImageDescriptor.camera_linear()

    def source_bounds(self):
        return self.image.metadata.active_area

    def raw_demosaic_identity(self):
        return self.demosaic

    def input_region(self, output: Rect, source_bounds: Rect) -> Rect:
        """RawEngine.cpp:290-297."""
        a = self.image.metadata.active_area
        if source_bounds != a:
            raise InvalidArgument("RAW source bounds differ from active area")
        radius = 6 if self.demosaic.algorithm == "rawengine.menon_base" and a.width > 1 and a.height > 1 else 1
        return _expand(output, a, radius)


class RasterSourceNode(Node):
    def __init__(self, image: RasterImage):
        self.image = image

    def output_descriptor(self):
        return ImageDescriptor.scene_linear(self.image.metadata.working_space)

    def supports_level(self, level):
        return (level.mip == 0 and level.quality in (RenderQuality.Final, RenderQuality.Preview)) or _reduced(level)

    def source_bounds(self):
        return Rect(0, 0, self.image.width, self.image.height)


class WhiteBalanceNode(Node):
    def __init__(self, input: Node, red: float, green: float, blue: float):
        if input is None:
            raise InvalidArgument("white balance input is null")
        if input.output_descriptor() != ImageDescriptor.camera_linear():
            raise InvalidArgument("white balance requires camera-linear RGB input")
        for g in (red, green, blue):
            g32 = np.float32(g)
            if not np.isfinite(g32) or g32 <= 0 or g32 > 65536:
                raise InvalidArgument("white balance gains must be finite and in (0, 65536]")
        self.input = input
        self.gains = (float(np.float32(red)), float(np.float32(green)), float(np.float32(blue)))

    def output_descriptor(self):
        return ImageDescriptor.camera_linear()

    def _ops(self):
        return [(C.OP_GAINS, self.gains)]


class ExposureNode(Node):
    def __init__(self, input: Node, stops: float):
        if input is None:
            raise InvalidArgument("exposure stops must be finite and in [-32, 32]")
        mult = ctypes.c_float()
        C.check(C.lib().rawops_exposure_multiplier(float(np.float32(stops)), ctypes.byref(mult)))
        d = input.output_descriptor()
        if d != ImageDescriptor.camera_linear() and d not in SCENE:
            raise InvalidArgument("exposure requires linear RGB input")
        self.input = input
        self.multiplier = mult.value
        self.descriptor = d

    def output_descriptor(self):
        return self.descriptor

    def supports_level(self, level):
        if _final(level):
            return True
        return _reduced(level) and self.descriptor.domain == "scene_linear" and self.input.supports_level(level)

    def _ops(self):
        return [(C.OP_SCALE, (self.multiplier,))]


class CameraToWorkingNode(Node):
    def __init__(self, input: Node, transform: CameraColorTransform, native_input_bounds: Rect = Rect()):
        if input is None or input.output_descriptor() != ImageDescriptor.camera_linear():
            raise InvalidArgument("color transform requires camera-linear RGB input")
        b = native_input_bounds
        if (b.width == 0) != (b.height == 0) or (not b.width and (b.x or b.y)):
            raise InvalidArgument("invalid camera preview input bounds")
        m = (ctypes.c_float * 9)()
        cam = (ctypes.c_double * 9)(*[float(v) for v in transform.camera_to_xyz_d50])
        C.check(C.lib().rawops_camera_to_working_matrix(cam, int(transform.target), m))
        self.input = input
        self.matrix = tuple(m)
        self.descriptor = ImageDescriptor.scene_linear(transform.target)
        self.native_bounds = b

    def output_descriptor(self):
        return self.descriptor

    def supports_level(self, level):
        if _final(level):
            return self.input.supports_level(level)
        return (_reduced(level) and self.native_bounds.width > 0 and self.native_bounds.height > 0
                and self.input.supports_level(RenderLevel()))

    def _ops(self):
        return [(C.OP_MATRIX, self.matrix)]


class WorkingSpaceConvertNode(Node):
    def __init__(self, input: Node, target: WorkingSpace):
        if input is None:
            raise InvalidArgument("working-space conversion input is null")
        d = input.output_descriptor()
        if d not in SCENE:
            raise InvalidArgument("working-space conversion requires declared scene-linear RGB")
        source = SCENE.index(d)
        self.input = input
        self.identity = source == int(target)
        self.descriptor = ImageDescriptor.scene_linear(target)
        if not self.identity:
            m = (ctypes.c_float * 9)()
            C.check(C.lib().rawops_working_to_working_matrix(source, int(target), m))
            self.matrix = tuple(m)

    def output_descriptor(self):
        return self.descriptor

    def supports_level(self, level):
        return _final(level) or (_reduced(level) and self.input.supports_level(level))

    def _ops(self):
        return [] if self.identity else [(C.OP_MATRIX, self.matrix)]


class WorkingToSrgbNode(Node):
    def __init__(self, input: Node):
        if input is None:
            raise InvalidArgument("working-to-sRGB input is null")
        d = input.output_descriptor()
        if d not in SCENE:
            raise InvalidArgument("sRGB conversion requires declared scene-linear working RGB")
        m = (ctypes.c_float * 9)()
        C.check(C.lib().rawops_working_to_srgb_matrix(SCENE.index(d), m))
        self.input = input
        self.matrix = tuple(m)

    def output_descriptor(self):
        return ImageDescriptor.linear_srgb()

    def supports_level(self, level):
        return _final(level) or (_reduced(level) and self.input.supports_level(level))

    def _ops(self):
        return [(C.OP_MATRIX, self.matrix)]


class ToneCurveNode(Node):
    def __init__(self, input: Node, shoulder: float, gamma: float):
        s32, g32 = np.float32(shoulder), np.float32(gamma)
        with np.errstate(all="ignore"):
            inv = np.float32(1.0) / g32
        if input is None or not np.isfinite(s32) or s32 <= 0 or not np.isfinite(g32) or g32 <= 0 or not np.isfinite(inv):
            raise InvalidArgument("tone shoulder and gamma must be finite and positive")
        src = input.output_descriptor()
        if src != ImageDescriptor.camera_linear() and src not in SCENE and src != ImageDescriptor.linear_srgb():
            raise InvalidArgument("tone curve requires linear RGB input")
        self.input = input
        self.shoulder, self.inverse_gamma = float(s32), float(inv)
        if src == ImageDescriptor.linear_srgb():
            self.descriptor = ImageDescriptor.display_linear_srgb()
        else:
            self.descriptor = ImageDescriptor.tone_mapped(src.primaries, src.white_point)

    def output_descriptor(self):
        return self.descriptor

    def supports_level(self, level):
        if _final(level):
            return True
        return (_reduced(level) and self.input.output_descriptor() == ImageDescriptor.linear_srgb()
                and self.input.supports_level(level))

    def _ops(self):
        return [(C.OP_TONE, (self.shoulder, self.inverse_gamma))]


class OutputClipNode(Node):
    def __init__(self, input: Node):
        if input is None or input.output_descriptor().domain != "tone_mapped":
            raise InvalidArgument("output clip requires tone-mapped RGB input")
        self.input = input
        d = input.output_descriptor()
        self.descriptor = ImageDescriptor("bounded", d.primaries, d.white_point, d.transfer)

    def output_descriptor(self):
        return self.descriptor

    def _ops(self):
        return [(C.OP_CLIP, ())]


class SrgbEncodeNode(Node):
    def __init__(self, input: Node):
        if input is None:
            raise InvalidArgument("sRGB encode input is null")
        if input.output_descriptor() != ImageDescriptor.display_linear_srgb():
            raise InvalidArgument("sRGB encode requires display-linear sRGB input")
        self.input = input

    def output_descriptor(self):
        return ImageDescriptor.srgb_output()

    def supports_level(self, level):
        return _final(level) or (_reduced(level) and self.input.supports_level(level))

    def _ops(self):
        return [(C.OP_SRGB_ENCODE, ())]


class ExtendedCurvesNode(Node):
    """ExtendedCurvesNode (ToneOps.hpp:33-50): master then channel, FP64."""

    def __init__(self, input: Node, settings: ExtendedCurvesSettings = None):
        settings = settings or ExtendedCurvesSettings()
        if input is None or input.output_descriptor() not in SCENE:
            raise InvalidArgument("extended curves/levels require scene-linear working RGB")
        self.input = input
        self.settings = settings
        knots = [list(settings.master)] + [list(c) for c in settings.channels]
        xs, ys, off = [], [], [0]
        for k in knots:
            for p in k:
                px, py = (p.x, p.y) if isinstance(p, CurvePoint) else p
                xs.append(float(px))
                ys.append(float(py))
            off.append(len(xs))
        self._xs = (ctypes.c_double * max(1, len(xs)))(*xs)
        self._ys = (ctypes.c_double * max(1, len(ys)))(*ys)
        self._off = off
        # Validate now (the reference validates in its constructor).
        Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], self._ops())

    def output_descriptor(self):
        return self.input.output_descriptor()

    def supports_level(self, level):
        return ((level.mip == 0 and level.quality in (RenderQuality.Final, RenderQuality.Preview))
                or _reduced(level)) and self.input.supports_level(level)

    def _ops(self):
        return [(C.OP_CURVES, (), (self._xs, self._ys, self._off,
                                   int(self.settings.interpolation == CurveInterpolation.ShapePreservingCubic)))]


def validate_working_y_guided_filter_settings(settings: WorkingYGuidedFilterSettings, max_radius=8):
    """SpatialOps.cpp:159-163 (radius 0..8 unless an extended radius is asked for)."""
    e = settings.epsilon
    if settings.radius > max_radius or not math.isfinite(e) or e < 2.0 ** -24 or e > 65536.0:
        raise InvalidArgument("guided filter requires radius 0..8 and finite epsilon [2^-24,65536]")


class WorkingYGuidedFilterNode(Node):
    """WorkingYGuidedFilterNode (SpatialOps.hpp:53-69).  ``max_radius`` may be
    raised to 32 to admit the radii beyond the reference's 0..8."""

    def __init__(self, input: Node, native_bounds: Rect, settings: WorkingYGuidedFilterSettings = None,
                 max_radius: int = 8):
        settings = settings or WorkingYGuidedFilterSettings()
        if not native_bounds.width or not native_bounds.height:
            raise InvalidArgument("spatial operation requires nonempty addressable bounds")
        validate_working_y_guided_filter_settings(settings, max_radius)
        if input is None:
            raise InvalidArgument("guided filter requires input")
        if input.output_descriptor() not in SCENE:
            raise InvalidArgument("guided filter requires scene-linear working RGB")
        self.input = input
        self.native_bounds = native_bounds
        self.settings = settings
        self.descriptor = input.output_descriptor()

    def output_descriptor(self):
        return self.descriptor

    def supports_level(self, level):
        return (((level.mip == 0 and level.quality in (RenderQuality.Final, RenderQuality.Preview)) or _reduced(level))
                and self.input.supports_level(level))

    def input_region(self, output: Rect, source_bounds: Rect) -> Rect:
        return _expand(output, source_bounds, 2 * self.settings.radius)


def _expand(r: Rect, area: Rect, radius: int) -> Rect:
    """expand_rect RawEngine.cpp:49-64 (validates containment first)."""
    _validate_rect(area, r)
    x0 = max(area.x, r.x - radius)
    y0 = max(area.y, r.y - radius)
    x1 = min(area.x + area.width, r.x + r.width + radius)
    y1 = min(area.y + area.height, r.y + r.height + radius)
    return Rect(x0, y0, x1 - x0, y1 - y0)


def _validate_rect(area: Rect, r: Rect):
    if (r.x < area.x or r.y < area.y or r.x + r.width > area.x + area.width
            or r.y + r.height > area.y + area.height):
        raise OutOfRange("viewport is outside the active RAW area")


# ------------------------------------------------------------ plan compiler --

class Program:
    """A compiled device plan (owns the rawops_plan handle)."""

    def __init__(self, handle, keep):
        self._handle = handle
        self._free = C.lib().rawops_plan_destroy     # bound now: usable during interpreter exit
        self._keep = keep

    @staticmethod
    def from_ops(demosaic, anchor, native_ops, reduced_ops):
        keep = []

        def descs(ops):
            arr = (C.OpDesc * max(1, len(ops)))()
            for i, op in enumerate(ops):
                kind, vals = op[0], op[1]
                arr[i].kind = kind
                for j, v in enumerate(vals):
                    arr[i].v[j] = v
                if kind == C.OP_CURVES:
                    xs, ys, off, cubic = op[2]
                    arr[i].knots_x = ctypes.cast(xs, ctypes.POINTER(ctypes.c_double))
                    arr[i].knots_y = ctypes.cast(ys, ctypes.POINTER(ctypes.c_double))
                    for j in range(5):
                        arr[i].knot_offsets[j] = off[j]
                    arr[i].cubic = cubic
                    keep.append((xs, ys))
            keep.append(arr)
            return arr

        d = C.PlanDesc()
        d.demosaic = demosaic
        d.mip_anchor = 1 if anchor else 0
        d.n_native = len(native_ops)
        d.n_reduced = len(reduced_ops)
        d.native_ops = ctypes.cast(descs(native_ops), ctypes.POINTER(C.OpDesc))
        d.reduced_ops = ctypes.cast(descs(reduced_ops), ctypes.POINTER(C.OpDesc))
        h = ctypes.c_void_p()
        C.check(C.lib().rawops_plan_create(ctypes.byref(d), ctypes.byref(h)))
        return Program(h, keep)

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._free(h)
            self._handle = None


@dataclass
class CompiledChain:
    kind: str                       # "raw" or "raster"
    source: object
    program: Program
    guided: Optional[WorkingYGuidedFilterNode] = None


def compile_chain(output: Node) -> CompiledChain:
    """Fuse the chain ending at ``output`` into one device plan."""
    chain = []
    n = output
    while n is not None:
        chain.append(n)
        n = n.input_node()
    chain.reverse()
    source, points = chain[0], chain[1:]
    if isinstance(source, RawUnpackNode):
        demosaic = C.DEMOSAIC_MENON if source.demosaic.algorithm == "rawengine.menon_base" else C.DEMOSAIC_BILINEAR
        native, reduced, anchor = [], [], False
        for node in points:
            if isinstance(node, WorkingYGuidedFilterNode):
                raise InvalidArgument("guided filter over a RAW chain: render the RAW chain to a raster first")
            ops = node._ops()
            if anchor:
                reduced += ops
            else:
                native += ops
            if isinstance(node, CameraToWorkingNode) and node.native_bounds.width:
                anchor = True
        return CompiledChain("raw", source, Program.from_ops(demosaic, anchor, native, reduced))
    if isinstance(source, RasterSourceNode):
        guided, post = None, []
        for node in points:
            if isinstance(node, WorkingYGuidedFilterNode):
                if guided is not None or post:
                    raise InvalidArgument("only one guided filter directly over the raster source is fused")
                guided = node
                continue
            post += node._ops()
        return CompiledChain("raster", source, Program.from_ops(C.DEMOSAIC_BILINEAR, False, [], post), guided)
    raise InvalidArgument("chain does not start at a RAW or raster source")


# ----------------------------------------------------------------- renderer --

class DeviceRender:
    """Device-side render entry: fills preallocated CUDA tensors on a stream."""

    @staticmethod
    def run(compiled: CompiledChain, viewport: Rect, level: RenderLevel, out=None, out_u8=None,
            flags=None, stream=None, workspace=None):
        t = _torch()
        if stream is None:
            stream = t.cuda.current_stream()
        sptr = ctypes.c_void_p(stream.cuda_stream)
        optr = ctypes.c_void_p(out.data_ptr()) if out is not None else None
        fptr = ctypes.c_void_p(flags.data_ptr()) if flags is not None else None
        stride = viewport.width
        if compiled.kind == "raw":
            u8 = ctypes.c_void_p(out_u8.data_ptr()) if out_u8 is not None else None
            C.check(C.lib().rawops_render(compiled.program.handle, compiled.source.image._handle, viewport.c(),
                                          level.mip, int(level.quality), optr, u8, stride, fptr, sptr))
        else:
            g = compiled.guided
            radius = g.settings.radius if g else 0
            eps = g.settings.epsilon if g else 1.0
            ws = ctypes.c_void_p(workspace.data_ptr()) if workspace is not None else None
            wsn = workspace.numel() * workspace.element_size() if workspace is not None else 0
            C.check(C.lib().rawops_guided_filter(compiled.source.image._handle, viewport.c(), level.mip, radius,
                                                 float(eps), compiled.program.handle, optr, stride, ws, wsn,
                                                 fptr, sptr))


class Renderer:
    """Renderer (RawEngine.hpp:661-710): render_tiles / render_image / render_roi."""

    TileCallback = Callable[[Tile], None]

    def _resolve(self, target, source_bounds):
        if isinstance(target, ImageGraph):
            return target.output(), target.source_bounds()
        return target, source_bounds

    def render_tiles(self, target, *args, cancellation: CancellationToken = None):
        """render_tiles(graph, request, callback) or (node, source_bounds, request, callback);
        legacy (…, viewport: Rect, callback, tile_size) forms are accepted too."""
        if isinstance(target, ImageGraph):
            node, bounds, rest = target.output(), target.source_bounds(), list(args)
        else:
            node, bounds, rest = target, args[0], list(args[1:])
        if isinstance(rest[0], Rect):
            viewport, callback = rest[0], rest[1]
            tile = rest[2] if len(rest) > 2 else 256
            request = RenderRequest(viewport, tile)
        else:
            request, callback = rest[0], rest[1]
        self._tiles(node, bounds, request, callback, cancellation)

    def _tiles(self, node, bounds, request, callback, cancellation):
        _validate_rect(request_bounds(bounds, request.level), request.viewport)
        if not node.supports_level(request.level):
            raise InvalidArgument("node does not support this render level")
        if callback is None or not request.tile_size:
            raise InvalidArgument("callback and tile size are required")
        if cancellation is not None and cancellation.is_cancelled():
            raise RenderCancelled("render cancelled")
        image = self._render_viewport(node, request.viewport, request.level)
        v, T = request.viewport, request.tile_size
        desc = node.output_descriptor()
        for y in range(v.y, v.y + v.height, T):
            for x in range(v.x, v.x + v.width, T):
                if cancellation is not None and cancellation.is_cancelled():
                    raise RenderCancelled("render cancelled")
                w, h = min(T, v.x + v.width - x), min(T, v.y + v.height - y)
                callback(Tile(Rect(x, y, w, h), image[y - v.y:y - v.y + h, x - v.x:x - v.x + w].copy(), desc))

    def _render_viewport_device(self, node, viewport, level):
        """The viewport rendered into a device tensor (h, w, 3), flags checked."""
        t = _torch()
        compiled = compile_chain(node)
        out = t.empty((viewport.height, viewport.width, 3), dtype=t.float32, device="cuda")
        flags = t.zeros(1, dtype=t.int32, device="cuda")
        if viewport.width and viewport.height:
            DeviceRender.run(compiled, viewport, level, out=out, flags=flags)
            C.check(C.lib().rawops_check_flags(ctypes.c_void_p(flags.data_ptr()),
                                               ctypes.c_void_p(t.cuda.current_stream().cuda_stream)))
        return out

    def _render_viewport(self, node, viewport, level):
        return self._render_viewport_device(node, viewport, level).cpu().numpy()

    def render_image(self, target, *args, cancellation: CancellationToken = None) -> Tile:
        if isinstance(target, ImageGraph):
            node, bounds, rest = target.output(), target.source_bounds(), list(args)
        else:
            node, bounds, rest = target, args[0], list(args[1:])
        request = rest[0] if isinstance(rest[0], RenderRequest) else RenderRequest(rest[0], rest[1] if len(rest) > 1 else 256)
        _validate_rect(request_bounds(bounds, request.level), request.viewport)
        if not node.supports_level(request.level):
            raise InvalidArgument("node does not support this render level")
        if cancellation is not None and cancellation.is_cancelled():
            raise RenderCancelled("render cancelled")
        img = self._render_viewport(node, request.viewport, request.level)
        return Tile(request.viewport, img, node.output_descriptor())

    def render_roi(self, graph, viewport: Rect, tile_size: int = 256, cancellation=None):
        return self.render_image(graph, viewport, tile_size, cancellation=cancellation).rgb.reshape(-1)


class ImageGraph:
    """ImageGraph (RawEngine.hpp:621-642; RawEngine.cpp:1306-1376).  The extra
    ``demosaic`` selects Menon for RAW input (the reference reaches Menon only
    through RawUnpackNode(image, identity)); ``linear_output`` ends the chain
    after calibration for FP32 scene-linear output."""

    def __init__(self, image, recipe: GraphRecipe = None, display_transform=None,
                 demosaic: RawDemosaicIdentity = RawDemosaicIdentity(), linear_output: bool = False):
        recipe = recipe or GraphRecipe()
        self.recipe = recipe
        if display_transform is not None or recipe.output_mode == OutputMode.IccDisplay:
            raise InvalidArgument("ICC display output is not supported by this library")
        if recipe.output_mode not in (OutputMode.LegacyBounded, OutputMode.SrgbPreview):
            raise InvalidArgument("unknown output mode")
        if isinstance(image, RawImage):
            self._raw = image
            self._bounds = image.metadata.active_area
            n = RawUnpackNode(image, demosaic)
            n = WhiteBalanceNode(n, recipe.red_gain, recipe.green_gain, recipe.blue_gain)
        elif isinstance(image, RasterImage):
            self._raw = None
            self._bounds = Rect(0, 0, image.width, image.height)
            if (recipe.red_gain, recipe.green_gain, recipe.blue_gain) != (1.0, 1.0, 1.0) or recipe.camera_color:
                raise InvalidArgument("raster input does not accept RAW calibration controls")
            n = RasterSourceNode(image)
        else:
            raise InvalidArgument("scene-linear source and valid bounds are required")
        n = ExposureNode(n, recipe.exposure_stops)
        if recipe.camera_color is not None and self._raw is not None:
            n = CameraToWorkingNode(n, recipe.camera_color, self._bounds)
        if linear_output:
            self._output = n
            return
        if recipe.output_mode == OutputMode.SrgbPreview:
            n = WorkingToSrgbNode(n)
        n = ToneCurveNode(n, recipe.tone_shoulder, recipe.tone_gamma)
        n = SrgbEncodeNode(n) if recipe.output_mode == OutputMode.SrgbPreview else OutputClipNode(n)
        self._output = n

    def image(self):
        if self._raw is None:
            raise RawOpsError("graph has a raster source, not a RAW source")
        return self._raw

    def source_bounds(self):
        return self._bounds

    def output(self):
        return self._output
