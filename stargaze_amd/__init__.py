"""MI355X-native replacement of stargaze's progressive path-tracer dispatch.

The compute lives in hand-written HIP kernels for gfx950 behind the C ABI in
``include/stargaze_amd.h``; this package holds the host-side mirror of the
reference's ``PathTracer`` interface and the reference byte layouts.
"""
from .layouts import GLOBALS, PRIMITIVE, RAY_STAR, SETTINGS, SPHERE  # noqa: F401
from .pathtracer import Frame, Options, PathTracer  # noqa: F401

__all__ = ["Frame", "Options", "PathTracer", "GLOBALS", "PRIMITIVE", "RAY_STAR", "SETTINGS", "SPHERE"]
