"""The reference's ``rawengine_native`` module (RawEnginePython.cpp) on MI355X:
``render``, ``render_raster``, ``RawSession``, ``RasterSession``,
``RenderJob`` and ``RenderCancelled`` with the reference's option keys, result
tuples and exceptions, rendering on the GPU through the C++ drop-in layer
(_rawengine_native, built in-tree).  Extension: ``render_dlpack`` returns the
result as a DLPack capsule of device memory for ``torch.from_dlpack``.

torch is imported first when it is installed: torch ships its own HIP runtime
under the same soname, and a process must use one HIP runtime only.
"""
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for this module
    pass

from ._rawengine_native import (RasterSession, RawSession, RenderCancelled, RenderJob,  # noqa: F401
                                create_icc_profile, icc_available, render, render_dlpack, render_raster)
