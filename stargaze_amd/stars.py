"""Catalogue stars in the tracer's ray format.

Restates the host side of stargaze's background star field: the seeded
catalogue itself (``stars::generate``, src/stars.rs:14-115), the radiance of
a catalogue star's drawn disc (``stars::radiance``, src/stars.rs:132-139) and
``stars::ray_catalogue`` (src/stars.rs:152-205), which packs the stars into
``RayStar`` records (unit direction, sin^2 of the angular radius, radiance)
and a conservative 256 x 128 spherical cell grid (32 769 offsets, then star
indices) so a miss ray tests only the discs overlapping its cell. The two
arrays are what ``sg_create`` / ``PathTracer(catalogue=, sky_cells=)`` take.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .layouts import RAY_STAR

GRID_W, GRID_H = 256, 128
POINT_SOURCE_GAIN = 100.0  # src/stars.rs:130
BRIGHTEST_STAR_IRRADIANCE = 6.4e-10 * POINT_SOURCE_GAIN  # src/stars.rs:121


@dataclass
class CatalogueStar:
    """``CatalogueStar``: direction, colour, relative brightness, drawn radius (rad)."""

    dir: tuple
    color: tuple = (1.0, 1.0, 1.0)
    bright: float = 1.0
    angular_radius: float = 1e-3


F32 = np.float32


class _Rng:
    """xorshift64* (src/stars.rs:14-32)."""

    def __init__(self, seed: int):
        self.state = (seed | 1) & 0xFFFFFFFFFFFFFFFF

    def u32(self) -> int:
        x = self.state
        x ^= x >> 12
        x ^= (x << 25) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 27
        self.state = x
        return ((x * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF) >> 32

    def f32(self) -> np.float32:
        return F32(self.u32()) / F32(4294967295.0)


def _srgb_to_linear(c: np.float32) -> np.float32:
    c = min(max(c, F32(0.0)), F32(1.0))
    return c / F32(12.92) if c <= F32(0.04045) else np.power((c + F32(0.055)) / F32(1.055), F32(2.4))


def blackbody_linear(temp_k) -> tuple:
    """Approximate blackbody colour (Tanner Helland's fit) in linear RGB,
    f32 (src/stars.rs:34-64)."""
    t = min(max(F32(temp_k) / F32(100.0), F32(10.0)), F32(400.0))
    r = F32(255.0) if t <= 66.0 else F32(329.69873) * np.power(t - F32(60.0), F32(-0.13320476))
    g = (F32(99.4708) * np.log(t) - F32(161.11957) if t <= 66.0
         else F32(288.12217) * np.power(t - F32(60.0), F32(-0.07551485)))
    b = (F32(255.0) if t >= 66.0 else F32(0.0) if t <= 19.0
         else F32(138.51773) * np.log(t - F32(10.0)) - F32(305.0448))
    return tuple(float(_srgb_to_linear(F32(c) / F32(255.0))) for c in (r, g, b))


def _normalize(v):
    x, y, z = (F32(c) for c in v)
    inv = F32(1.0) / np.sqrt(x * x + y * y + z * z)
    return x * inv, y * inv, z * inv


def _random_dir(rng: _Rng) -> tuple:
    z = F32(1.0) - F32(2.0) * rng.f32()
    phi = F32(2.0 * math.pi) * rng.f32()
    r = np.sqrt(max(F32(1.0) - z * z, F32(0.0)))
    return float(r * np.cos(phi)), float(r * np.sin(phi)), float(z)


def _star(rng: _Rng, dim: float) -> CatalogueStar:
    u = rng.f32()
    bright = (F32(0.02) + u * u * u * F32(1.6)) * F32(dim)
    x = rng.f32()
    temp = F32(2500.0) + F32(12000.0) * (x * x)
    radius = min(max(F32(0.7) + F32(2.6) * np.sqrt(bright), F32(0.7)), F32(3.4)) * F32(0.0004)
    return CatalogueStar(dir=_random_dir(rng), color=blackbody_linear(temp), bright=float(bright),
                         angular_radius=float(radius))


def generate(count: int, band: int, seed: int) -> list:
    """``count`` stars uniform over the sky, then ``band`` dimmer, cooler stars
    bulging about a great circle (the Milky Way), from one seeded stream
    (src/stars.rs:91-115). The circle's in-plane axes come from the
    branch-free orthonormal basis of Duff et al. (2017), as glam's
    ``any_orthonormal_vector`` builds them."""
    rng = _Rng(seed)
    out = [_star(rng, 1.0) for _ in range(count)]
    n = _normalize((0.35, 0.82, 0.45))
    sign = F32(1.0) if n[2] >= 0.0 else F32(-1.0)
    a = F32(-1.0) / (sign + n[2])
    b = n[0] * n[1] * a
    u = (b, sign + n[1] * n[1] * a, -n[1])
    v = _normalize((n[1] * u[2] - u[1] * n[2], n[2] * u[0] - u[2] * n[0], n[0] * u[1] - u[0] * n[1]))
    for _ in range(band):
        angle = F32(2.0 * math.pi) * rng.f32()
        spread = (rng.f32() + rng.f32() + rng.f32() - F32(1.5)) * F32(0.28)
        c, s_ = np.cos(angle), np.sin(angle)
        d = _normalize(tuple(u[i] * c + v[i] * s_ + n[i] * spread for i in range(3)))
        star = _star(rng, 0.28)
        star.dir = tuple(float(x) for x in d)
        star.color = blackbody_linear(F32(2500.0) + F32(6000.0) * rng.f32())
        out.append(star)
    return out


def radiance(star: CatalogueStar) -> np.ndarray:
    """Irradiance of the star spread over its drawn disc (src/stars.rs:132-139)."""
    radius = np.float32(min(max(float(star.angular_radius), 1e-6), 0.1))
    irradiance = np.float32(BRIGHTEST_STAR_IRRADIANCE) * np.float32(star.bright) / np.float32(1.62)
    value = irradiance / (np.float32(math.pi) * radius * radius)
    return np.asarray(star.color, np.float32) * value


def ray_catalogue(stars) -> tuple[np.ndarray, np.ndarray]:
    """``(RayStar records, cell offsets + indices)`` (src/stars.rs:152-205)."""
    bins = [[] for _ in range(GRID_W * GRID_H)]
    data = np.zeros(max(len(stars), 1), RAY_STAR)
    for index, star in enumerate(stars):
        d = np.asarray(star.dir, np.float64)
        d = d / np.linalg.norm(d)
        radius = min(max(float(star.angular_radius), 1e-6), 0.1)
        data["direction"][index] = [d[0], d[1], d[2], math.sin(radius) ** 2]
        data["radiance"][index] = [*radiance(star), 0.0]
        theta = math.acos(max(-1.0, min(1.0, d[2])))
        phi = math.atan2(d[1], d[0]) + math.pi
        y0 = min(int(max(theta - radius, 0.0) / math.pi * GRID_H), GRID_H - 1)
        y1 = min(int(min(theta + radius, math.pi) / math.pi * GRID_H), GRID_H - 1)
        if theta <= radius or theta + radius >= math.pi:
            lon = math.pi
        else:
            lon = math.asin(max(0.0, min(1.0, math.sin(radius) / math.sin(theta))))
        for x in range(GRID_W):
            cell_phi = (x + 0.5) * 2.0 * math.pi / GRID_W
            separation = abs(((cell_phi - phi + math.pi) % (2.0 * math.pi)) - math.pi)
            if separation <= lon + math.pi / GRID_W + 1e-6:
                for y in range(y0, y1 + 1):
                    bins[y * GRID_W + x].append(index)
    offsets, indices = [], []
    for b in bins:
        offsets.append(len(indices))
        indices.extend(b)
    offsets.append(len(indices))
    cells = np.array(offsets + indices, np.uint32)
    return (data if len(stars) else data[:0]), cells
