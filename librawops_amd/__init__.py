"""MI355X-native (gfx950) tiled RAW render path with the RawEngine render API.

The compute runs in hand-written HIP kernels (librawops_hip.so, C ABI in
include/rawops_hip.h); this package is the Python host mirror of the
reference's ImageGraph / RenderRequest / Renderer API plus the
differentiable and batched forms.  There is no CPU execution path.
"""
from .graph import (BayerPattern, CameraColorTransform, CancellationToken, CurveInterpolation,  # noqa: F401
                    CurvePoint, DomainError, ExposureNode, ExtendedCurvesNode, ExtendedCurvesSettings,
                    FilterOverflow, GraphRecipe, ImageDescriptor, ImageGraph, InvalidArgument, Node,
                    OutOfRange, OutputClipNode, OutputMode, RasterImage, RasterMetadata, RasterSourceNode,
                    RawDemosaicIdentity, RawImage, RawMetadata, RawOpsError, RawUnpackNode,
                    Rect, RenderCancelled, RenderLevel, RenderQuality, RenderRequest, Renderer,
                    SrgbEncodeNode, Tile, ToneCurveNode, WhiteBalanceNode, WorkingSpace,
                    WorkingSpaceConvertNode, WorkingToSrgbNode, WorkingYGuidedFilterNode,
                    WorkingYGuidedFilterSettings, CameraToWorkingNode, compile_chain, request_bounds,
                    validate_raw_demosaic, validate_working_y_guided_filter_settings)

from .analysis import ChannelHistogram, HistogramOptions, RgbHistogram, histogram_rgb  # noqa: F401

MENON = RawDemosaicIdentity("rawengine.menon_base", 1)
BILINEAR = RawDemosaicIdentity("rawengine.bilinear", 1)
