"""ORACLE — test infrastructure only.

ctypes front end of the C restatement (``sg_oracle.c``) plus a Python
restatement of the host logic of ``PathTracer::encode``
(src/pathtracer.rs:164-188 History, :789-1009 encode). Tests, the smoke check
and the bench's CPU-baseline leg use it as the checker; the product never
imports it.

The layouts below are declared independently of the product package, so a
layout mistake on either side shows up as a parity failure.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"

_f4 = ("<f4", (4,))
_u4 = ("<u4", (4,))
GLOBALS = np.dtype(
    [(n, *_f4) for n in ("cam_right", "cam_up", "cam_forward", "viewport", "atmo_center", "atmo_rayleigh",
                         "atmo_params", "ground_east", "ground_up", "ground_north")]
    + [("ground_counts", *_u4)]
)
SETTINGS = np.dtype([("globals", GLOBALS), ("counts", *_u4), ("sampling", *_u4), ("background", *_f4), ("lights", *_u4)])
PROBE_RAY = np.dtype([("offset", "<f4", (3,)), ("anchor", "<u4"), ("direction", "<f4", (3,)), ("target", "<u4")])
PROBE_RESULT = np.dtype(
    [("distance", "<f8"), ("index", "<u4"), ("pad", "<u4"), ("normal", "<f8", (3,)), ("extra", "<f8", (3,))]
)
BODY_BYTES, PRIM_BYTES = 160, 48
assert GLOBALS.itemsize == 176 and SETTINGS.itemsize == 240
assert PROBE_RAY.itemsize == 32 and PROBE_RESULT.itemsize == 64

MAX_BATCH_SAMPLES = 64  # src/pathtracer.rs:22
MAX_SAMPLES = 1 << 24  # src/graphics.rs:11


class OrcScene(ctypes.Structure):
    _fields_ = [
        ("settings", ctypes.c_void_p),
        ("bodies", ctypes.c_void_p),
        ("objects", ctypes.c_void_p),
        ("sky_light", ctypes.c_void_p),
        ("terrain", ctypes.c_void_p),
        ("catalogue", ctypes.c_void_p),
        ("catalogue_count", ctypes.c_uint32),
        ("sky_cells", ctypes.c_void_p),
        ("milky_way", ctypes.c_void_p),
        ("ring_profile", ctypes.c_void_p),
        ("ring_profile_body", ctypes.c_uint32),
        ("air_view_steps", ctypes.c_uint32),
        ("air_sun_steps", ctypes.c_uint32),
        ("mie_height", ctypes.c_float),
    ]


def build() -> None:
    """Compile both oracle builds (gcc) into oracle/_build/."""
    subprocess.run(["make", "-C", os.fspath(HERE), "-s"], check=True)


_libs: dict[str, ctypes.CDLL] = {}


def library(precision: str = "f32") -> ctypes.CDLL:
    if precision in _libs:
        return _libs[precision]
    path = BUILD / f"liborc_{precision}.so"
    if not path.exists():
        build()
    lib = ctypes.CDLL(os.fspath(path))
    P, U32 = ctypes.c_void_p, ctypes.c_uint32
    lib.orc_hash.argtypes, lib.orc_hash.restype = [U32], U32
    lib.orc_random.argtypes, lib.orc_random.restype = [ctypes.POINTER(U32)], ctypes.c_double
    lib.orc_encode_rows.argtypes, lib.orc_encode_rows.restype = [ctypes.POINTER(OrcScene), U32, U32, P, P, P], None
    lib.orc_terrain_cache.argtypes, lib.orc_terrain_cache.restype = [ctypes.POINTER(OrcScene)], None
    lib.orc_sky_cache.argtypes, lib.orc_sky_cache.restype = [ctypes.POINTER(OrcScene), P], None
    lib.orc_probe.argtypes, lib.orc_probe.restype = [ctypes.POINTER(OrcScene), U32, P, U32, P], None
    lib.orc_ring_energy.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, U32]
    lib.orc_ring_energy.restype = ctypes.c_double
    lib.orc_real_bytes.argtypes, lib.orc_real_bytes.restype = [], U32
    lib.orc_set_counting.argtypes, lib.orc_set_counting.restype = [ctypes.c_int], None
    lib.orc_counters.argtypes, lib.orc_counters.restype = [P], None
    lib.orc_counter_count.argtypes, lib.orc_counter_count.restype = [], U32
    lib.orc_set_planetshine_thinning.argtypes, lib.orc_set_planetshine_thinning.restype = [ctypes.c_int], None
    _libs[precision] = lib
    return lib


def terrain_layout(cells: int) -> dict:
    """Float offsets of the terrain buffer (src/shaders/terrain.wgsl:6-20)."""
    width = cells + 1
    compact = cells > 1024
    vertices = width * width
    leaves = cells // 2
    nodes = (4 * leaves * leaves - 1) // 3
    bounds = (1 if compact else 3) * vertices
    cache = bounds + 2 * nodes
    stride = 2 if compact else 6
    return dict(cells=cells, width=width, compact=compact, vertices=vertices, nodes=nodes, bounds=bounds,
                cache=cache, vertex_cache=cache + 6 * nodes, stride=stride,
                total=cache + 6 * nodes + stride * vertices)


def _check_ground(frame) -> None:
    """The library's frame check on the ground theme: the basalt ground (4)
    has no restated materials here or in the library, and themes past the
    dunes (5) do not exist."""
    g = np.asarray(frame.globals)
    theme = int(g["ground_counts"][2]) & 255
    if g["ground_north"][3] != 0 and (theme == 4 or theme > 5):
        raise ValueError(f"ground theme {theme} is not built")


class OracleTracer:
    """``PathTracer`` host logic over the C restatement, on the CPU."""

    def __init__(self, dimensions, options=None, precision="f32", catalogue=None, sky_cells=None, milky_way=None):
        self.lib = library(precision)
        self.planetshine_thinning = True
        self.width, self.height = int(dimensions[0]), int(dimensions[1])
        o = options
        self.noise_seed = int(getattr(o, "noise_seed", 0))
        self.samples_per_frame = int(getattr(o, "samples_per_frame", 1))
        self.max_bounces = int(getattr(o, "max_bounces", 8))
        self.sample_limit = int(getattr(o, "sample_limit", MAX_SAMPLES))
        self.exposure = float(getattr(o, "exposure", 1.0))
        self.milky_way_gain = float(getattr(o, "milky_way", 4.0e-4))
        self.key = b""
        self.history_samples = 0
        self.sequence = 0
        self.catalogue = catalogue
        self.sky_cells = sky_cells
        self.milky_way = milky_way
        self.terrain = None
        self.terrain_cells = 0
        self.terrain_radius = None
        self.sky = np.zeros((128, 4), np.float32)
        self.ring_profile = None
        self.air = (8, 4, 0.0)
        self.moments_on = False
        self._alloc()

    def _alloc(self):
        n = self.width * self.height
        self.accumulation = np.zeros((n, 4), np.float32)
        self.backdrop = np.zeros((n, 4), np.float32)
        self.moments = np.zeros((n, 4), np.float32)

    def samples(self) -> int:
        return self.history_samples

    def set_planetshine_thinning(self, on: bool) -> None:
        """Test knob: with ``on`` False this tracer's passes never thin
        planetshine (orc_set_planetshine_thinning in sg_oracle.c)."""
        self.planetshine_thinning = bool(on)

    def enable_moments(self, on=True):
        self.moments_on = on

    def set_ring_profile(self, body: int, tau):
        """``PathTracer.set_ring_profile``: 512 bins, or None; restarts."""
        self.ring_profile = None if tau is None else (int(body), np.array(tau, np.float32).reshape(512))
        self.key = b""
        self.history_samples = 0

    def set_air_model(self, view_steps=8, light_steps=4, mie_scale_height=0.0):
        """Air march and separate Mie scale height (AU); restarts."""
        self.air = (int(view_steps), int(light_steps), float(mie_scale_height))
        self.key = b""
        self.history_samples = 0

    def upload_terrain(self, data: np.ndarray, cells: int):
        lay = terrain_layout(cells)
        assert data.size == lay["cache"], "terrain data size does not match its cell count"
        t = np.zeros(lay["total"], np.float32)
        t[: data.size] = data
        self.terrain = t
        self.terrain_cells = cells
        self.terrain_radius = None

    def set_quality(self, samples, bounces, limit):
        bounces = min(max(bounces, 1), 64)
        if self.max_bounces != bounces:
            self.key = b""
            self.history_samples = 0
        self.samples_per_frame = min(max(samples, 1), 64)
        self.max_bounces = bounces
        self.sample_limit = min(max(limit, 1), MAX_SAMPLES)

    def resize(self, dimensions):
        if (self.width, self.height) == tuple(dimensions):
            return
        self.width, self.height = int(dimensions[0]), int(dimensions[1])
        self._alloc()
        self.key = b""
        self.history_samples = 0

    def _history(self, frame):
        key = (np.asarray(frame.globals).tobytes() + np.asarray(frame.bodies).tobytes()
               + np.asarray(frame.ground).tobytes() + np.float64(frame.scene_time).tobytes()
               + np.uint32(self.max_bounces).tobytes() + np.float32(self.milky_way_gain).tobytes())
        if key != self.key:
            self.key = key
            self.history_samples = 0

    def settings(self, frame, count):
        """Settings packing of src/pathtracer.rs:803-859."""
        st = np.zeros((), SETTINGS)
        g = np.frombuffer(np.asarray(frame.globals).tobytes(), GLOBALS)[0].copy()
        vp = g["viewport"].copy()
        vp[3] = np.float32(vp[0]) / np.float32(max(vp[1], np.float32(1.0)))
        vp[0], vp[1], vp[2] = self.width, self.height, self.exposure
        g["viewport"] = vp
        top = np.float32(0.0)
        ground = np.frombuffer(np.asarray(frame.ground).tobytes(), np.float32).reshape(-1, 12)
        for p in ground:
            e = np.float32(np.sqrt(np.float32(p[4] * p[4] + p[5] * p[5] + p[6] * p[6]), dtype=np.float32))
            top = max(top, np.float32(p[1] + e))
        gc = g["ground_counts"].copy()
        gc[3] = np.float32(top).view(np.uint32)
        cells = self.terrain_cells or 512
        gc[2] = int(gc[2]) | (int(cells).bit_length() - 1) << 8
        g["ground_counts"] = gc
        st["globals"] = g
        bodies = np.frombuffer(np.asarray(frame.bodies).tobytes(), np.float32).reshape(-1, 40)
        lights = [0, 0, 0, 0]
        for i, b in enumerate(bodies):
            if b[7] >= 0.5:
                if lights[3] < 3:
                    lights[lights[3]] = i
                lights[3] += 1
        st["lights"] = lights
        st["sampling"] = [(self.noise_seed + self.sequence) & 0xFFFFFFFF, 0, 0, 0]
        st["counts"] = [len(bodies), self.history_samples, count, self.max_bounces]
        mw = self.milky_way
        st["background"] = [self.milky_way_gain, mw.shape[1] if mw is not None else 1024,
                            mw.shape[0] if mw is not None else 512, 100.0]
        return st

    def _scene(self, st, frame):
        keep = [st]
        bodies = np.ascontiguousarray(np.frombuffer(np.asarray(frame.bodies).tobytes() or b"\0" * BODY_BYTES, np.uint8))
        objects = np.ascontiguousarray(np.frombuffer(np.asarray(frame.ground).tobytes() or b"\0" * PRIM_BYTES, np.uint8))
        keep += [bodies, objects]
        sc = OrcScene()
        sc.settings = st.ctypes.data
        sc.bodies = bodies.ctypes.data
        sc.objects = objects.ctypes.data
        sc.sky_light = self.sky.ctypes.data
        sc.air_view_steps, sc.air_sun_steps, sc.mie_height = self.air
        sc.terrain = self.terrain.ctypes.data if self.terrain is not None else None
        if self.catalogue is not None and len(self.catalogue):
            cat = np.ascontiguousarray(self.catalogue)
            cells = np.ascontiguousarray(self.sky_cells, np.uint32)
            keep += [cat, cells]
            sc.catalogue = cat.ctypes.data
            sc.catalogue_count = len(cat)
            sc.sky_cells = cells.ctypes.data
        if self.milky_way is not None:
            mw = np.ascontiguousarray(self.milky_way, np.uint32)
            keep.append(mw)
            sc.milky_way = mw.ctypes.data
        if self.ring_profile is not None:
            body, tau = self.ring_profile
            if body < len(np.asarray(frame.bodies).reshape(-1)):
                t = np.ascontiguousarray(tau, np.float32)
                keep.append(t)
                sc.ring_profile = t.ctypes.data
                sc.ring_profile_body = body
        return sc, keep

    def encode(self, frame, rows=None):
        """One progressive pass; ``rows`` limits the rows traced (sampling only)."""
        _check_ground(frame)
        if np.asarray(frame.globals)["ground_counts"][1] != 0 and self.terrain is None:
            # The library refuses the same frame (sg_encode, SG_E_STATE).
            raise RuntimeError("terrain enabled but no terrain uploaded (upload_terrain)")
        self._history(frame)
        count = min(self.samples_per_frame, MAX_BATCH_SAMPLES, max(self.sample_limit - self.history_samples, 0))
        st = self.settings(frame, count)
        if count > 0:
            sc, keep = self._scene(st, frame)
            self.lib.orc_set_planetshine_thinning(int(self.planetshine_thinning))
            g = st["globals"]
            if g["ground_counts"][1] != 0 and g["ground_north"][3] != 0:
                radius = float(g["ground_up"][3])
                if self.terrain_radius != radius:
                    self.lib.orc_terrain_cache(ctypes.byref(sc))
                    self.terrain_radius = radius
            if self.history_samples == 0:
                self.lib.orc_sky_cache(ctypes.byref(sc), self.sky.ctypes.data)
            r0, r1 = (0, self.height) if rows is None else rows
            self.lib.orc_encode_rows(ctypes.byref(sc), r0, r1, self.accumulation.ctypes.data,
                                     self.backdrop.ctypes.data, self.moments.ctypes.data if self.moments_on else None)
        self.history_samples += count
        self.sequence = (self.sequence + count) & 0xFFFFFFFF
        return count

    def read_accumulation(self):
        return self.accumulation.reshape(self.height, self.width, 4).copy()

    def read_backdrop(self):
        return self.backdrop.reshape(self.height, self.width, 4).copy()

    def read_moments(self):
        return self.moments.reshape(self.height, self.width, 4).copy()

    def probe(self, frame, kind, rays):
        _check_ground(frame)
        if (kind == 4 or np.asarray(frame.globals)["ground_counts"][1] != 0) and self.terrain is None:
            # Kind 4 is the terrain query; the library refuses the same probe
            # (sg_probe, SG_E_STATE).
            raise RuntimeError("terrain probe without uploaded terrain")
        st = self.settings(frame, 0)
        sc, keep = self._scene(st, frame)
        g = st["globals"]
        if self.terrain is not None:
            radius = float(g["ground_up"][3])
            if self.terrain_radius != radius:
                self.lib.orc_terrain_cache(ctypes.byref(sc))
                self.terrain_radius = radius
        r = np.ascontiguousarray(np.frombuffer(np.asarray(rays).tobytes(), PROBE_RAY))
        out = np.zeros(r.size, PROBE_RESULT)
        self.lib.orc_probe(ctypes.byref(sc), int(kind), r.ctypes.data, r.size, out.ctypes.data)
        return out


EVENTS = ("paths", "segments", "sphere", "sphere_hit", "ring", "light_sample", "visibility", "ring_optics",
          "atmo_cell", "density_step", "vnoise", "box", "rock", "vegetation", "terrain_node", "triangle",
          "vertex", "rng")


def count_events(tracer: "OracleTracer", frame, rows, stages: bool = False):
    """Event counts of one pass over ``rows`` (the work model): all of them,
    or with ``stages`` a pair (everything but the camera-air step, the
    camera-air step)."""
    lib = tracer.lib
    assert lib.orc_counter_count() == len(EVENTS)
    lib.orc_set_counting(1)
    try:
        tracer.encode(frame, rows=rows)
        out = np.zeros((2, len(EVENTS)), np.uint64)
        lib.orc_counters(out.ctypes.data)
    finally:
        lib.orc_set_counting(0)
    parts = [dict(zip(EVENTS, (int(v) for v in row))) for row in out]
    if stages:
        return parts[0], parts[1]
    return {k: parts[0][k] + parts[1][k] for k in EVENTS}
