"""ctypes binding of ``libstargaze_amd.so`` (the HIP library).

The library is built in-tree by ``stargaze_amd.build`` and loaded from the
package directory. There is no fallback: if the library is missing or fails
to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libstargaze_amd.so"

SG_OK = 0
ERRORS = {-1: "invalid argument", -2: "capacity exceeded", -3: "device error", -4: "invalid state"}

# Every symbol include/stargaze_amd.h declares.
EXPORTS = (
    "sg_create",
    "sg_destroy",
    "sg_set_milky_way",
    "sg_resize",
    "sg_samples",
    "sg_size",
    "sg_needs_redraw",
    "sg_prepare_frame",
    "sg_set_quality",
    "sg_set_exposure_controls",
    "sg_upload_terrain",
    "sg_encode",
    "sg_read_accumulation",
    "sg_read_backdrop",
    "sg_enable_moments",
    "sg_read_moments",
    "sg_write_accumulation",
    "sg_write_backdrop",
    "sg_read_exposure",
    "sg_display",
    "sg_display_srgb8",
    "sg_display_srgb8_noalloc",
    "sg_synchronize",
    "sg_set_profiling",
    "sg_kernel_time",
    "sg_step_time",
    "sg_set_integrator",
    "sg_set_tail",
    "sg_get_launch",
    "sg_set_air_queue",
    "sg_get_air_queue",
    "sg_autotune",
    "sg_last_error",
    "sg_probe",
    "sg_set_ring_profile",
    "sg_set_density_profile",
    "sg_restart",
    "sg_pass_id",
    "sg_set_air_model",
    "sg_get_air_model",
    "sg_adjoint",
    "sg_enable_adjoint",
    "sg_adjoint_time",
    "sg_batch_create",
    "sg_batch_destroy",
    "sg_batch_set_launch",
    "sg_batch_get_launch",
    "sg_batch_encode",
    "sg_batch_synchronize",
    "sg_batch_view_times",
    "sg_batch_set_profiling",
    "sg_batch_kernel_time",
    "sg_batch_encode_views",
    "sg_batch_autotune",
    "sg_batch_last_error",
)


class SgFrame(ctypes.Structure):
    """``sg_frame`` (include/stargaze_amd.h)."""

    _fields_ = [
        ("globals", ctypes.c_void_p),
        ("bodies", ctypes.c_void_p),
        ("body_count", ctypes.c_uint32),
        ("ground", ctypes.c_void_p),
        ("ground_count", ctypes.c_uint32),
        ("scene_time", ctypes.c_double),
    ]


class SgLaunch(ctypes.Structure):
    """``sg_launch`` (include/stargaze_amd.h)."""

    _fields_ = [("integrator", ctypes.c_uint32), ("chunk_paths", ctypes.c_uint32), ("tail_paths", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class SgBatchLaunch(ctypes.Structure):
    """``sg_batch_launch`` (include/stargaze_amd.h)."""

    _fields_ = [
        ("block", ctypes.c_uint32),
        ("workgroups", ctypes.c_uint32),
        ("blocks_per_pull", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class SgAirModel(ctypes.Structure):
    """``sg_air_model`` (include/stargaze_amd.h)."""

    _fields_ = [
        ("view_steps", ctypes.c_uint32),
        ("light_steps", ctypes.c_uint32),
        ("mie_scale_height", ctypes.c_float),
        ("reserved", ctypes.c_uint32),
    ]


class NativeError(RuntimeError):
    pass


_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load the HIP library, failing loudly when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise NativeError(
            f"{LIBRARY} is missing: build it with `python -m stargaze_amd.build` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback"
        )
    lib = ctypes.CDLL(os.fspath(LIBRARY))
    P, U32, I32, U64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    sig = {
        "sg_create": ([ctypes.POINTER(P), U32, U32, P, U32, P, U32, P], I32),
        "sg_destroy": ([P], I32),
        "sg_set_milky_way": ([P, P, U32, U32], I32),
        "sg_resize": ([P, U32, U32], I32),
        "sg_samples": ([P], U32),
        "sg_size": ([P, ctypes.POINTER(U32), ctypes.POINTER(U32)], I32),
        "sg_needs_redraw": ([P], I32),
        "sg_prepare_frame": ([P, ctypes.POINTER(SgFrame)], I32),
        "sg_set_quality": ([P, U32, U32, U32], I32),
        "sg_set_exposure_controls": ([P, I32, ctypes.c_float], I32),
        "sg_upload_terrain": ([P, P, U64, U32], I32),
        "sg_encode": ([P, ctypes.POINTER(SgFrame)], I32),
        "sg_read_accumulation": ([P, P, U64], I32),
        "sg_read_backdrop": ([P, P, U64], I32),
        "sg_enable_moments": ([P, I32], I32),
        "sg_read_moments": ([P, P, U64], I32),
        "sg_write_accumulation": ([P, P, U64], I32),
        "sg_write_backdrop": ([P, P, U64], I32),
        "sg_read_exposure": ([P, P, P], I32),
        "sg_display": ([P, U32, U32, P], I32),
        "sg_display_srgb8": ([P, U32, U32, P], I32),
        "sg_display_srgb8_noalloc": ([P, U32, U32, P], I32),
        "sg_synchronize": ([P], I32),
        "sg_set_profiling": ([P, I32], I32),
        "sg_step_time": ([P, U32, P, P], I32),
        "sg_kernel_time": ([P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(U64)], I32),
        "sg_set_integrator": ([P, U32, U32], I32),
        "sg_set_tail": ([P, U32], I32),
        "sg_get_launch": ([P, ctypes.POINTER(SgLaunch)], I32),
        "sg_set_air_queue": ([P, U32], I32),
        "sg_get_air_queue": ([P, ctypes.POINTER(U32)], I32),
        "sg_autotune": ([P, ctypes.POINTER(SgFrame), U32, ctypes.POINTER(SgLaunch), ctypes.POINTER(ctypes.c_double)],
                        I32),
        "sg_last_error": ([P], ctypes.c_char_p),
        "sg_probe": ([P, ctypes.POINTER(SgFrame), U32, P, U32, P], I32),
        "sg_set_ring_profile": ([P, U32, P, U32], I32),
        "sg_restart": ([P], I32),
        "sg_pass_id": ([P], U64),
        "sg_set_air_model": ([P, P], I32),
        "sg_get_air_model": ([P, P], I32),
        "sg_set_density_profile": ([P, P, U32], I32),
        "sg_adjoint": ([P, ctypes.POINTER(SgFrame), P, U64, P, U32, P], I32),
        "sg_adjoint_time": ([P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(U64)], I32),
        "sg_enable_adjoint": ([P, I32], I32),
        "sg_batch_create": ([ctypes.POINTER(P), ctypes.POINTER(P), U32], I32),
        "sg_batch_destroy": ([P], I32),
        "sg_batch_set_launch": ([P, ctypes.POINTER(SgBatchLaunch)], I32),
        "sg_batch_get_launch": ([P, ctypes.POINTER(SgBatchLaunch)], I32),
        "sg_batch_encode": ([P, P], I32),
        "sg_batch_synchronize": ([P], I32),
        "sg_batch_view_times": ([P, P, U32, ctypes.POINTER(ctypes.c_double)], I32),
        "sg_batch_set_profiling": ([P, I32], I32),
        "sg_batch_kernel_time": ([P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(U64)], I32),
        "sg_batch_encode_views": ([P, P, P, I32], I32),
        "sg_batch_autotune": ([P, P, U32, ctypes.POINTER(SgBatchLaunch), ctypes.POINTER(ctypes.c_double)], I32),
        "sg_batch_last_error": ([P], ctypes.c_char_p),
    }
    for name in EXPORTS:
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = sig[name]
    _lib = lib
    return lib


def check(lib: ctypes.CDLL, ctx, rc: int, what: str, batch: bool = False) -> int:
    if rc < 0:
        last = lib.sg_batch_last_error if batch else lib.sg_last_error
        detail = last(ctx).decode() if ctx else ""
        raise NativeError(f"{what} failed: {ERRORS.get(rc, rc)}: {detail}")
    return rc
