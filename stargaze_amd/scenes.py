"""Scene builders: the reference's test frames and the benchmark workloads.

Frames follow the reference's conventions (src/main.rs:822-1018): telescope
space looks along -Z, lengths in AU, body centres split into an f32 high part
and an f32 residual, star radiance ``2.5 * luminosity / radius_au**2``
(src/main.rs:71,846). Workloads are synthetic, seeded and self-contained.
"""
from __future__ import annotations

import math

import numpy as np

from .layouts import AU_KM, GLOBALS, PRIMITIVE, SPHERE
from .pathtracer import Frame

METRES_PER_AU = AU_KM * 1000.0
SUN_LIGHT = 2.5  # src/main.rs:71


def metres(x: float) -> float:
    """Metres to AU."""
    return x / METRES_PER_AU


# -- the reference's unit-test helpers (src/pathtracer_tests.rs:32-61,
#    src/ring_tests.rs:4-10) -------------------------------------------------
def frame() -> Frame:
    g = np.zeros((), GLOBALS)
    g["cam_right"] = [1, 0, 0, 0]
    g["cam_up"] = [0, 1, 0, 0]
    g["cam_forward"] = [0, 0, -1, 1e-7]
    g["viewport"] = [1, 1, 1, 0]
    return Frame(globals=g)


def body(center, radius, color, emissive: bool) -> np.ndarray:
    b = np.zeros((), SPHERE)
    b["center"] = center
    b["radius"] = radius
    b["color"] = color
    b["emissive"] = 1.0 if emissive else 0.0
    return b


def ringed(center, radius, pole, tau) -> np.ndarray:
    planet = body(center, radius, [0.0] * 3, False)
    planet["ring_plane"] = [pole[0], pole[1], pole[2], 1.25]
    planet["ring_params"] = [2.4, tau, 0.0, 0.0]
    planet["ring_color"] = [0.8, 0.8, 0.8, 0.0]
    return planet


def bodies(*records) -> np.ndarray:
    return np.array([np.asarray(r, SPHERE) for r in records], SPHERE).reshape(-1)


# -- helpers for physical scenes ---------------------------------------------
def split(center_au: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """High/low f32 split of an f64 telescope-space centre (src/main.rs:838-844)."""
    c = np.asarray(center_au, np.float64)
    hi = c.astype(np.float32)
    lo = (c - hi.astype(np.float64)).astype(np.float32)
    return hi, lo


def place(b: np.ndarray, center_au) -> np.ndarray:
    hi, lo = split(center_au)
    b["center"] = hi
    b["center_low"][:3] = lo
    return b


def star(center_au, radius_au: float, luminosity: float = 1.0) -> np.ndarray:
    s = body([0, 0, 0], radius_au, [SUN_LIGHT * luminosity / radius_au**2] * 3, True)
    return place(s, center_au)


def direction(azimuth_deg: float, elevation_deg: float) -> np.ndarray:
    """Unit vector in a frame with +Y up and -Z forward (azimuth to +X)."""
    a, e = math.radians(azimuth_deg), math.radians(elevation_deg)
    return np.array([math.cos(e) * math.sin(a), math.sin(e), -math.cos(e) * math.cos(a)])


def camera(g: np.ndarray, fov_deg: float, width: int, height: int) -> None:
    g["cam_right"] = [1, 0, 0, 0]
    g["cam_up"] = [0, 1, 0, 0]
    g["cam_forward"] = [0, 0, -1, math.tan(math.radians(fov_deg) * 0.5)]
    g["viewport"] = [width, height, 1.0, 0.0]


class Pcg:
    """Seeded PCG32 for workload synthesis (same step as the path RNG)."""

    def __init__(self, seed: int):
        self.state = seed & 0xFFFFFFFF

    def next_u32(self) -> int:
        self.state = (self.state * 747796405 + 2891336453) & 0xFFFFFFFF
        s = self.state
        word = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
        return ((word >> 22) ^ word) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u32() >> 8) / 16777216.0


# -- workloads ---------------------------------------------------------------
def airless_triple(width: int = 256, height: int = 256) -> Frame:
    """Star + planet + moon, camera 2 m above the planet, 60 deg FOV.

    Star R = 7e8 m at 1.5e11 m; planet R = 6.4e6 m, albedo 0.3; moon
    R = 1.7e6 m, albedo 0.12, 3.8e8 m from the planet. Constant albedos
    (procedural strength 0), no rings, air, ground or catalogue.
    """
    f = frame()
    camera(f.globals, 60.0, width, height)
    r_planet = metres(6.4e6)
    up = np.array([0.0, 1.0, 0.0])
    planet_c = -up * (r_planet + metres(2.0))
    sun_dir = direction(-70.0, 25.0)
    moon_dir = direction(12.0, 9.0)
    moon_c = planet_c + moon_dir * metres(3.8e8)
    planet = place(body([0, 0, 0], r_planet, [0.3] * 3, False), planet_c)
    moon = place(body([0, 0, 0], metres(1.7e6), [0.12] * 3, False), moon_c)
    sun = star(sun_dir * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, planet, moon)
    return f


def ringed_giant(width: int = 1920, height: int = 1080, seed: int = 0x5EED1234) -> Frame:
    """Banded gas giant with a C/B/A ring system, four moons, one star.

    Giant R = 6.0e7 m, gas-giant bands between albedo 0.2 and 0.6; moons with
    radii log-uniform in [2e5, 2.6e6] m and albedo U[0.05, 0.8] (PCG seed
    0x5EED1234); equatorial ring 6.7e7-1.4e8 m with the banded optical-depth
    profile (C ~0.1, B ~1.5, A ~0.5, Cassini/Encke/Keeler gaps), single
    scattering albedo 0.6.
    """
    rng = Pcg(seed)
    f = frame()
    camera(f.globals, 50.0, width, height)
    r_giant = metres(6.0e7)
    giant_c = np.array([0.0, 0.0, -metres(3.6e8)])
    giant = body([0, 0, 0], r_giant, [0.6, 0.55, 0.45], False)
    giant["axis_x"] = [1, 0, 0, 1.0]  # style 1: gas giant
    pole = np.array([0.0, math.cos(math.radians(22.0)), math.sin(math.radians(22.0))])
    x_axis = np.array([1.0, 0.0, 0.0])
    y_axis = np.cross(pole, x_axis)
    giant["axis_x"][:3] = x_axis
    giant["axis_y"] = [*y_axis, 7.0]
    giant["surface_color"] = [0.2, 0.17, 0.13, 0.8]
    giant["surface_params"] = [7.0, 2.0, 0.0, 0.0]
    giant["ring_plane"] = [*pole, 6.7e7 / 6.0e7]
    giant["ring_params"] = [1.4e8 / 6.0e7, 1.0, 1.0, 0.0]
    giant["ring_color"] = [0.6, 0.6, 0.6, 0.0]
    place(giant, giant_c)
    moons = []
    for k in range(4):
        radius = metres(math.exp(rng.uniform(math.log(2e5), math.log(2.6e6))))
        albedo = rng.uniform(0.05, 0.8)
        angle = rng.uniform(0.0, 2 * math.pi)
        dist = metres(rng.uniform(1.6e8, 2.6e8))
        lift = rng.uniform(-0.08, 0.08)
        offset = np.array([math.cos(angle) * dist, lift * dist, math.sin(angle) * dist * 0.6])
        m = place(body([0, 0, 0], radius, [albedo] * 3, False), giant_c + offset)
        moons.append(m)
    sun = star(direction(-50.0, 30.0) * 9.5, metres(7.0e8))
    f.bodies = bodies(sun, giant, *moons)
    return f


WORKLOADS = {
    "airless_triple_256": (airless_triple, dict(width=256, height=256), 16, 4),
    "ringed_giant_1080p": (ringed_giant, dict(width=1920, height=1080), 64, 6),
}
