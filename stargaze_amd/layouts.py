"""Byte layouts shared with the reference's Rust host and WGSL shaders.

Each numpy dtype is byte-identical to the reference struct it names, so a
buffer built here can be handed to the C ABI (``include/stargaze_amd.h``)
unchanged:

* ``GLOBALS``   — ``Globals``  (src/renderer.rs:13-30), 176 B
* ``SPHERE``    — ``Sphere``   (src/renderer.rs:32-48) / WGSL ``Body``, 160 B
* ``PRIMITIVE`` — ``Primitive`` (src/ground.rs:164-170), 48 B
* ``RAY_STAR``  — ``RayStar``  (src/stars.rs:143-148), 32 B
* ``SETTINGS``  — ``Settings`` (src/pathtracer.rs:44-52), 240 B
"""
from __future__ import annotations

import numpy as np

F4 = ("<f4", (4,))
U4 = ("<u4", (4,))

GLOBALS = np.dtype(
    [
        ("cam_right", *F4),
        ("cam_up", *F4),
        ("cam_forward", *F4),
        ("viewport", *F4),
        ("atmo_center", *F4),
        ("atmo_rayleigh", *F4),
        ("atmo_params", *F4),
        ("ground_east", *F4),
        ("ground_up", *F4),
        ("ground_north", *F4),
        ("ground_counts", *U4),
    ]
)

SPHERE = np.dtype(
    [
        ("center", "<f4", (3,)),
        ("radius", "<f4"),
        ("color", "<f4", (3,)),
        ("emissive", "<f4"),
        ("center_low", *F4),
        ("ring_plane", *F4),
        ("ring_params", *F4),
        ("ring_color", *F4),
        ("axis_x", *F4),
        ("axis_y", *F4),
        ("surface_color", *F4),
        ("surface_params", *F4),
    ]
)

PRIMITIVE = np.dtype([("center", *F4), ("extent", *F4), ("albedo", *F4)])

RAY_STAR = np.dtype([("direction", *F4), ("radiance", *F4)])

SETTINGS = np.dtype(
    [
        ("globals", GLOBALS),
        ("counts", *U4),
        ("sampling", *U4),
        ("background", *F4),
        ("lights", *U4),
    ]
)

CAMERA_AIR = np.dtype([("radiance", "<f4", (3,)), ("rng", "<u4")])

EXPOSURE = np.dtype([("value", "<f4"), ("key", "<f4"), ("percentile", "<f4"), ("enabled", "<f4"), ("bias", "<f4")])
HISTOGRAM_BINS = 256  # src/pathtracer.rs:36
PROBE_RAY = np.dtype(
    [("offset", "<f4", (3,)), ("anchor", "<u4"), ("direction", "<f4", (3,)), ("target", "<u4")]
)

PROBE_RESULT = np.dtype(
    [
        ("distance", "<f4"),
        ("index", "<u4"),
        ("normal", "<f4", (3,)),
        ("extra", "<f4", (3,)),
    ]
)

OPTIONS = np.dtype(
    [
        ("noise_seed", "<u4"),
        ("samples_per_frame", "<u4"),
        ("max_bounces", "<u4"),
        ("sample_limit", "<u4"),
        ("exposure", "<f4"),
        ("auto_exposure", "<i4"),
        ("ev_bias", "<f4"),
        ("auto_key", "<f4"),
        ("auto_percentile", "<f4"),
        ("milky_way", "<f4"),
    ]
)

assert GLOBALS.itemsize == 176
assert SPHERE.itemsize == 160
assert PRIMITIVE.itemsize == 48
assert RAY_STAR.itemsize == 32
assert SETTINGS.itemsize == 240
assert CAMERA_AIR.itemsize == 16
assert PROBE_RAY.itemsize == 32
assert PROBE_RESULT.itemsize == 32
assert OPTIONS.itemsize == 40

MISS = 0xFFFFFFFF
LOCAL_BASE = 0x80000000  # src/shaders/ground.wgsl:14
TERRAIN_ID = 0x80100000  # src/shaders/terrain.wgsl:23

BODY_CAPACITY = 256  # src/pathtracer.rs:21
MAX_BATCH_SAMPLES = 64  # src/pathtracer.rs:22
GROUND_CAPACITY = 256  # src/ground.rs:8
SKY_TEXELS = 128  # src/ground.rs:9
MAX_SAMPLES = 1 << 24  # src/graphics.rs:11
POINT_SOURCE_GAIN = 100.0  # src/stars.rs:130
SKY_WIDTH, SKY_HEIGHT = 1024, 512  # src/sky.rs:8-9
AU_KM = 149_597_870.7  # src/sim.rs:13
METRES_PER_AU = 149_597_870_700.0  # src/shaders/ground.wgsl:15

# Differentiable rendering (include/stargaze_amd.h SG_GRAD_*).
RING_PROFILE_BINS = 512
DENSITY_BINS = 16
GRAD_RING = 3 * BODY_CAPACITY
GRAD_AIR = GRAD_RING + RING_PROFILE_BINS  # Rayleigh rgb, Mie, Mie g
GRAD_DENSITY = GRAD_AIR + 5
GRAD_COUNT = GRAD_DENSITY + DENSITY_BINS


def grad_albedo(body: int, channel: int) -> int:
    return 3 * body + channel


def zeros(dtype: np.dtype, n: int | None = None) -> np.ndarray:
    """A zero-filled record (``n is None``) or array of ``dtype``."""
    return np.zeros(() if n is None else n, dtype=dtype)
