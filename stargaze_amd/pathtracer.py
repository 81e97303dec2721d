"""Host-side mirror of the reference's ``PathTracer`` (src/pathtracer.rs:190-1015).

Same method names, argument meaning and error behaviour (errors raise where
the reference panics). All work runs in the HIP library through the C ABI;
this module only packs arrays and calls it.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .layouts import EXPOSURE, GLOBALS, GRAD_COUNT, HISTOGRAM_BINS, OPTIONS, PRIMITIVE, PROBE_RAY, PROBE_RESULT, RAY_STAR, SPHERE, MAX_SAMPLES


@dataclass
class Options:
    """``Options`` (src/pathtracer.rs:76-112), with the reference defaults."""

    noise_seed: int = 0
    samples_per_frame: int = 1
    max_bounces: int = 8
    sample_limit: int = MAX_SAMPLES
    exposure: float = 1.0
    auto_exposure: bool = False
    ev_bias: float = 0.0
    auto_key: float = 0.25
    auto_percentile: float = 0.9
    milky_way: float = 4.0e-4

    def pack(self) -> np.ndarray:
        o = np.zeros((), OPTIONS)
        for name in OPTIONS.names:
            o[name] = getattr(self, name)
        return o


@dataclass
class Frame:
    """``Frame`` (src/renderer.rs:50-56) without UI labels."""

    globals: np.ndarray  # GLOBALS record
    bodies: np.ndarray = field(default_factory=lambda: np.zeros(0, SPHERE))
    ground: np.ndarray = field(default_factory=lambda: np.zeros(0, PRIMITIVE))
    scene_time: float = 0.0

    def native(self) -> tuple[_native.SgFrame, tuple]:
        g = np.ascontiguousarray(np.asarray(self.globals, GLOBALS).reshape(()))
        b = np.ascontiguousarray(np.asarray(self.bodies, SPHERE).reshape(-1))
        p = np.ascontiguousarray(np.asarray(self.ground, PRIMITIVE).reshape(-1))
        f = _native.SgFrame(
            g.ctypes.data,
            b.ctypes.data if b.size else None,
            b.size,
            p.ctypes.data if p.size else None,
            p.size,
            float(self.scene_time),
        )
        return f, (g, b, p)  # keep the arrays alive for the call


class PathTracer:
    """One progressive accumulation on the current HIP device."""

    def __init__(
        self,
        dimensions: tuple[int, int],
        catalogue: np.ndarray | None = None,
        sky_cells: np.ndarray | None = None,
        options: Options | None = None,
    ):
        self._lib = _native.load()
        self.options = options or Options()
        ctx = ctypes.c_void_p()
        cat = np.ascontiguousarray(catalogue if catalogue is not None else np.zeros(0, RAY_STAR), RAY_STAR)
        cells = np.ascontiguousarray(sky_cells if sky_cells is not None else np.zeros(0, np.uint32), np.uint32)
        opt = self.options.pack()
        rc = self._lib.sg_create(
            ctypes.byref(ctx),
            int(dimensions[0]),
            int(dimensions[1]),
            cat.ctypes.data if cat.size else None,
            cat.size,
            cells.ctypes.data if cells.size else None,
            cells.size,
            opt.ctypes.data,
        )
        _native.check(self._lib, None, rc, "sg_create")
        self._ctx = ctx

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.sg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> int:
        return _native.check(self._lib, self._ctx, rc, what)

    # -- reference API ----------------------------------------------------
    def resize(self, dimensions: tuple[int, int]) -> None:
        self._check(self._lib.sg_resize(self._ctx, int(dimensions[0]), int(dimensions[1])), "resize")

    def samples(self) -> int:
        return int(self._lib.sg_samples(self._ctx))

    def size(self) -> tuple[int, int]:
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.sg_size(self._ctx, ctypes.byref(w), ctypes.byref(h)), "size")
        return int(w.value), int(h.value)

    def needs_redraw(self) -> bool:
        return bool(self._check(self._lib.sg_needs_redraw(self._ctx), "needs_redraw"))

    def prepare_frame(self, frame: Frame) -> bool:
        f, keep = frame.native()
        return bool(self._check(self._lib.sg_prepare_frame(self._ctx, ctypes.byref(f)), "prepare_frame"))

    def set_quality(self, samples: int, bounces: int, limit: int) -> None:
        self._check(self._lib.sg_set_quality(self._ctx, samples, bounces, limit), "set_quality")
        self.options.samples_per_frame = max(1, min(64, samples))
        self.options.max_bounces = max(1, min(64, bounces))
        self.options.sample_limit = max(1, min(MAX_SAMPLES, limit))

    def set_exposure_controls(self, auto: bool, bias: float) -> None:
        self._check(self._lib.sg_set_exposure_controls(self._ctx, int(auto), float(bias)), "set_exposure_controls")

    def encode(self, frame: Frame) -> None:
        f, keep = frame.native()
        self._check(self._lib.sg_encode(self._ctx, ctypes.byref(f)), "encode")

    # -- data -------------------------------------------------------------
    def upload_terrain(self, data: np.ndarray, cells: int) -> None:
        d = np.ascontiguousarray(data, np.float32)
        self._check(self._lib.sg_upload_terrain(self._ctx, d.ctypes.data, d.size, int(cells)), "upload_terrain")

    def set_milky_way(self, texels: np.ndarray) -> None:
        t = np.ascontiguousarray(texels, np.uint32)
        self._check(self._lib.sg_set_milky_way(self._ctx, t.ctypes.data, t.shape[1], t.shape[0]), "set_milky_way")

    def enable_moments(self, enable: bool = True) -> None:
        self._check(self._lib.sg_enable_moments(self._ctx, int(enable)), "enable_moments")

    def _read(self, fn, what: str) -> np.ndarray:
        w, h = self.size()
        out = np.empty((h, w, 4), np.float32)
        self._check(fn(self._ctx, out.ctypes.data, w * h), what)
        return out

    def read_accumulation(self) -> np.ndarray:
        """Linear HDR mean per pixel, (H, W, 4) f32 (rgb, backdrop luminance)."""
        return self._read(self._lib.sg_read_accumulation, "read_accumulation")

    def read_backdrop(self) -> np.ndarray:
        return self._read(self._lib.sg_read_backdrop, "read_backdrop")

    def read_moments(self) -> np.ndarray:
        return self._read(self._lib.sg_read_moments, "read_moments")

    def write_accumulation(self, rgba: np.ndarray) -> None:
        """Restore the accumulation image (H, W, 4); the history is unchanged."""
        self._write(self._lib.sg_write_accumulation, rgba, "write_accumulation")

    def write_backdrop(self, rgba: np.ndarray) -> None:
        self._write(self._lib.sg_write_backdrop, rgba, "write_backdrop")

    def _write(self, fn, rgba, what):
        w, h = self.size()
        a = np.ascontiguousarray(rgba, np.float32)
        if a.shape != (h, w, 4):
            raise ValueError(f"image must be ({h}, {w}, 4), got {a.shape}")
        self._check(fn(self._ctx, a.ctypes.data, w * h), what)

    def read_exposure(self) -> tuple[np.ndarray, np.ndarray]:
        """The display exposure state (EXPOSURE record) and the last metered
        luminance histogram (256 counts)."""
        state = np.zeros((), EXPOSURE)
        hist = np.zeros(HISTOGRAM_BINS, np.uint32)
        self._check(self._lib.sg_read_exposure(self._ctx, state.ctypes.data, hist.ctypes.data), "read_exposure")
        return state, hist

    def display(self, width: int, height: int) -> np.ndarray:
        """``PathTracer::display``: exposed, ACES-mapped (H, W, 4) f32 at the given size."""
        out = np.empty((height, width, 4), np.float32)
        self._check(self._lib.sg_display(self._ctx, int(width), int(height), out.ctypes.data), "display")
        return out

    def display_srgb8(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.uint8)
        self._check(self._lib.sg_display_srgb8(self._ctx, int(width), int(height), out.ctypes.data), "display_srgb8")
        return out

    def display_srgb8_noalloc(self, width: int, height: int, out: np.ndarray | None = None) -> np.ndarray:
        """``display_srgb8`` through the library's static per-GPU staging
        (no device allocation; at most 3840 x 2160 pixels)."""
        if out is None:
            out = np.empty((height, width, 4), np.uint8)
        if out.shape != (height, width, 4) or out.dtype != np.uint8 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous (height, width, 4) uint8 array")
        self._check(self._lib.sg_display_srgb8_noalloc(self._ctx, int(width), int(height), out.ctypes.data),
                    "display_srgb8_noalloc")
        return out

    def synchronize(self) -> None:
        self._check(self._lib.sg_synchronize(self._ctx), "synchronize")

    def set_profiling(self, enable: bool) -> None:
        self._check(self._lib.sg_set_profiling(self._ctx, int(enable)), "set_profiling")

    def kernel_time(self) -> tuple[float, int]:
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        self._check(self._lib.sg_kernel_time(self._ctx, ctypes.byref(ms), ctypes.byref(n)), "kernel_time")
        return float(ms.value), int(n.value)

    def set_integrator(self, integrator: int, chunk_paths: int = 0) -> None:
        """INTEGRATOR_AUTO (default), INTEGRATOR_FUSED or INTEGRATOR_WAVEFRONT
        (see include/stargaze_amd.h)."""
        self._check(self._lib.sg_set_integrator(self._ctx, int(integrator), int(chunk_paths)), "set_integrator")

    def restart(self) -> None:
        """Start a fresh accumulation with the next pass (new draws)."""
        self._check(self._lib.sg_restart(self._ctx), "restart")

    def set_ring_profile(self, body: int, tau: np.ndarray | None) -> None:
        """Tabulated ring optical depth (RING_PROFILE_BINS bins) for ``body``;
        ``None`` restores the banded profile."""
        if tau is None:
            self._check(self._lib.sg_set_ring_profile(self._ctx, int(body), None, 0), "set_ring_profile")
            return
        t = np.ascontiguousarray(tau, np.float32).reshape(-1)
        self._check(self._lib.sg_set_ring_profile(self._ctx, int(body), t.ctypes.data, t.size), "set_ring_profile")

    def set_air_model(self, view_steps: int = 8, light_steps: int = 4, mie_scale_height: float = 0.0) -> None:
        """Air march (view cells per camera ray, density samples per light
        path) and a separate Mie scale height in AU (0: one scale height for
        both species, as the reference's shader). Restarts the accumulation."""
        m = _native.SgAirModel(int(view_steps), int(light_steps), float(mie_scale_height), 0)
        self._check(self._lib.sg_set_air_model(self._ctx, ctypes.byref(m)), "set_air_model")

    @property
    def air_model(self) -> dict:
        m = _native.SgAirModel()
        self._check(self._lib.sg_get_air_model(self._ctx, ctypes.byref(m)), "get_air_model")
        return {"view_steps": int(m.view_steps), "light_steps": int(m.light_steps),
                "mie_scale_height": float(m.mie_scale_height)}

    def pass_id(self) -> int:
        """Identity of the last pass that rendered samples (changes with every
        such pass and every restart)."""
        return int(self._lib.sg_pass_id(self._ctx))

    def set_density_profile(self, density: np.ndarray | None) -> None:
        """Relative air density per altitude slice (DENSITY_BINS); ``None``
        restores exp(-h / H)."""
        if density is None:
            self._check(self._lib.sg_set_density_profile(self._ctx, None, 0), "set_density_profile")
            return
        d = np.ascontiguousarray(density, np.float32).reshape(-1)
        self._check(self._lib.sg_set_density_profile(self._ctx, d.ctypes.data, d.size), "set_density_profile")

    def adjoint(self, frame: Frame, d_image: np.ndarray, want_mean: bool = False):
        """Gradient (GRAD_COUNT floats) of sum(d_image * pass_mean.rgb) for the
        last encoded pass of ``frame``; with ``want_mean`` also the pass mean
        (H, W, 4) the adjoint recomputed."""
        w, h = self.size()
        d = np.ascontiguousarray(d_image, np.float32)
        if d.shape != (h, w, 3):
            raise ValueError(f"d_image must be ({h}, {w}, 3), got {d.shape}")
        grad = np.zeros(GRAD_COUNT, np.float32)
        mean = np.zeros((h, w, 4), np.float32) if want_mean else None
        f, keep = frame.native()
        self._check(
            self._lib.sg_adjoint(self._ctx, ctypes.byref(f), d.ctypes.data, w * h, grad.ctypes.data, GRAD_COUNT,
                                 mean.ctypes.data if want_mean else None),
            "adjoint",
        )
        return (grad, mean) if want_mean else grad

    def enable_adjoint(self, enable: bool = True) -> None:
        """Keep per-sample primal outputs so ``adjoint`` replays each path once."""
        self._check(self._lib.sg_enable_adjoint(self._ctx, int(enable)), "enable_adjoint")

    def step_time(self, step: int) -> tuple[float, int]:
        """Device ms and launches of one wavefront step (STEP_EXTEND,
        STEP_CAMERA_AIR, STEP_SHADE, STEP_AIR_OCCLUSION) since set_profiling."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        self._check(self._lib.sg_step_time(self._ctx, int(step), ctypes.byref(ms), ctypes.byref(n)), "step_time")
        return float(ms.value), int(n.value)

    def adjoint_time(self) -> tuple[float, int]:
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        self._check(self._lib.sg_adjoint_time(self._ctx, ctypes.byref(ms), ctypes.byref(n)), "adjoint_time")
        return float(ms.value), int(n.value)

    def set_tail(self, tail_paths: int) -> None:
        """Wavefront compaction threshold (0: compact to the end)."""
        self._check(self._lib.sg_set_tail(self._ctx, int(tail_paths)), "set_tail")

    @property
    def air_queue(self) -> int:
        """Records per path of the deferred camera-air shadow-ray queue."""
        n = ctypes.c_uint32()
        self._check(self._lib.sg_get_air_queue(self._ctx, ctypes.byref(n)), "get_air_queue")
        return int(n.value)

    def set_air_queue(self, records_per_path: int) -> None:
        """Size (records per path, 1-256) of the deferred camera-air
        shadow-ray queue; a chunk that overflows it grows it and runs again."""
        self._check(self._lib.sg_set_air_queue(self._ctx, int(records_per_path)), "set_air_queue")

    @property
    def launch(self) -> dict:
        l = _native.SgLaunch()
        self._check(self._lib.sg_get_launch(self._ctx, ctypes.byref(l)), "get_launch")
        return l.as_dict()

    def autotune(self, frame: Frame, repeats: int = 1) -> tuple[dict, float]:
        """Fastest integrator, queue capacity and compaction threshold for
        ``frame``; the accumulation restarts with the next pass."""
        f, keep = frame.native()
        best = _native.SgLaunch()
        ms = ctypes.c_double()
        self._check(self._lib.sg_autotune(self._ctx, ctypes.byref(f), int(repeats), ctypes.byref(best),
                                          ctypes.byref(ms)), "autotune")
        return best.as_dict(), float(ms.value)

    def probe(self, frame: Frame, kind: int, rays: np.ndarray) -> np.ndarray:
        """Run production device helpers on ``rays`` (PROBE_RAY records)."""
        r = np.ascontiguousarray(rays, PROBE_RAY).reshape(-1)
        out = np.zeros(r.size, PROBE_RESULT)
        f, keep = frame.native()
        self._check(
            self._lib.sg_probe(self._ctx, ctypes.byref(f), int(kind), r.ctypes.data, r.size, out.ctypes.data), "probe"
        )
        return out


INTEGRATOR_AUTO, INTEGRATOR_FUSED, INTEGRATOR_WAVEFRONT = 0, 1, 2
STEP_EXTEND, STEP_CAMERA_AIR, STEP_SHADE, STEP_AIR_OCCLUSION = range(4)
(PROBE_SPHERE, PROBE_CLOSEST, PROBE_RING, PROBE_VISIBILITY, PROBE_TERRAIN, PROBE_LANDSCAPE, PROBE_CONE, PROBE_ALBEDO,
 PROBE_GROUND, PROBE_NOISE, PROBE_SKY) = range(11)

BATCH_ORDER_GIVEN, BATCH_ORDER_SHORTEST, BATCH_ORDER_LONGEST = 0, 1, 2


class Batch:
    """Many views rendered one progressive pass at a time in one launch.

    No reference counterpart (the reference has one ``PathTracer`` per view,
    src/pathtracer.rs:190); each view is a ``PathTracer`` whose images are
    read as usual. The batch must be closed before its views.
    """

    def __init__(self, views: list[PathTracer]):
        self._lib = _native.load()
        self.views = list(views)
        handles = (ctypes.c_void_p * len(self.views))(*[v._ctx.value for v in self.views])
        b = ctypes.c_void_p()
        rc = self._lib.sg_batch_create(ctypes.byref(b), handles, len(self.views))
        _native.check(self._lib, None, rc, "sg_batch_create", batch=True)
        self._b = b

    def close(self) -> None:
        if getattr(self, "_b", None):
            self._lib.sg_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> int:
        return _native.check(self._lib, self._b, rc, what, batch=True)

    def _frames(self, frames: list[Frame]):
        if len(frames) != len(self.views):
            raise ValueError(f"{len(frames)} frames for {len(self.views)} views")
        arr = (_native.SgFrame * len(frames))()
        keep = []
        for i, f in enumerate(frames):
            native, k = f.native()
            arr[i] = native
            keep.append(k)
        return arr, keep

    @property
    def launch(self) -> dict:
        l = _native.SgBatchLaunch()
        self._check(self._lib.sg_batch_get_launch(self._b, ctypes.byref(l)), "sg_batch_get_launch")
        return l.as_dict()

    def set_launch(self, **kw) -> None:
        cur = self.launch
        cur.update(kw)
        l = _native.SgBatchLaunch(**cur)
        self._check(self._lib.sg_batch_set_launch(self._b, ctypes.byref(l)), "sg_batch_set_launch")

    def encode(self, frames: list[Frame]) -> None:
        arr, keep = self._frames(frames)
        self._check(self._lib.sg_batch_encode(self._b, arr), "sg_batch_encode")

    def encode_views(self, frames: list[Frame], serial: bool = False) -> np.ndarray:
        """The per-view form: one ``encode`` per view, overlapping on the
        device or (``serial``) one after another as on one queue; returns
        each view's completion time (ms) after the first launch was queued."""
        arr, keep = self._frames(frames)
        ms = np.zeros(len(self.views), np.float64)
        self._check(self._lib.sg_batch_encode_views(self._b, arr, ms.ctypes.data, int(serial)),
                    "sg_batch_encode_views")
        return ms

    def synchronize(self) -> None:
        self._check(self._lib.sg_batch_synchronize(self._b), "sg_batch_synchronize")

    def view_times(self) -> tuple[np.ndarray, float]:
        """Completion time (ms) of each view in the last batch pass and the pass span."""
        ms = np.zeros(len(self.views), np.float64)
        span = ctypes.c_double()
        self._check(
            self._lib.sg_batch_view_times(self._b, ms.ctypes.data, len(self.views), ctypes.byref(span)),
            "sg_batch_view_times",
        )
        return ms, float(span.value)

    def set_profiling(self, enable: bool) -> None:
        self._check(self._lib.sg_batch_set_profiling(self._b, int(enable)), "sg_batch_set_profiling")

    def kernel_time(self) -> tuple[float, int]:
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        self._check(self._lib.sg_batch_kernel_time(self._b, ctypes.byref(ms), ctypes.byref(n)), "sg_batch_kernel_time")
        return float(ms.value), int(n.value)

    def autotune(self, frames: list[Frame], repeats: int = 2) -> tuple[dict, float]:
        """Pick the fastest launch settings; resets every view's history."""
        arr, keep = self._frames(frames)
        best = _native.SgBatchLaunch()
        ms = ctypes.c_double()
        self._check(
            self._lib.sg_batch_autotune(self._b, arr, int(repeats), ctypes.byref(best), ctypes.byref(ms)),
            "sg_batch_autotune",
        )
        return best.as_dict(), float(ms.value)
