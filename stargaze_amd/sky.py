"""The Milky Way panorama the tracer samples on camera escape rays.

Restates ``sky::panorama`` (src/sky.rs:8-117): a 1024 x 512 map of linear
RGB10 texels (one u32 each, r | g << 10 | b << 20) of a procedural galactic
band (value-noise clumps, wisps, rifts and a central bulge) in galactic
longitude and latitude. The reference embeds the map in the binary at
binding 6; here the caller generates it once and hands it over with
``PathTracer.set_milky_way`` (``sg_set_milky_way``). Evaluated in float32
with numpy, so texels may differ from the reference's by the last unit of
the 10-bit quantisation where a value sits on a rounding edge.
"""
from __future__ import annotations

import functools

import numpy as np

WIDTH, HEIGHT = 1024, 512
_f = np.float32


def _lattice(x, y, z):
    """src/sky.rs:11-19: hashed lattice value in [0, 1) (wrapping u32 products)."""
    with np.errstate(over="ignore"):
        return _lattice_u32(x, y, z)


def _lattice_u32(x, y, z):
    h = ((x.astype(np.uint32) * np.uint32(0x8DA6B343))
         ^ (y.astype(np.uint32) * np.uint32(0xD8163841))
         ^ (z.astype(np.uint32) * np.uint32(0xCB1AB31F))
         ^ np.uint32(0x5EED1234))
    h = (h ^ (h >> np.uint32(16))) * np.uint32(0x7FEB352D)
    h = (h ^ (h >> np.uint32(15))) * np.uint32(0x846CA68B)
    return ((h ^ (h >> np.uint32(16))) >> np.uint32(8)).astype(_f) / _f(16_777_216.0)


def _noise(p):
    """src/sky.rs:21-42: quintic-faded trilinear value noise; p (..., 3)."""
    floor = np.floor(p)
    cell = floor.astype(np.int64).astype(np.int32)
    t = (p - floor).astype(_f)
    t = t * t * t * (t * (t * _f(6.0) - _f(15.0)) + _f(10.0))
    cx, cy, cz = cell[..., 0], cell[..., 1], cell[..., 2]
    tx, ty, tz = t[..., 0], t[..., 1], t[..., 2]

    def mix(a, b, s):
        return a + (b - a) * s

    def plane(z):
        return mix(mix(_lattice(cx, cy, z), _lattice(cx + 1, cy, z), tx),
                   mix(_lattice(cx, cy + 1, z), _lattice(cx + 1, cy + 1, z), tx), ty)

    return mix(plane(cz), plane(cz + 1), tz)


def _clouds(p):
    """src/sky.rs:44-53: four octaves of lacunarity 2.07."""
    total = np.zeros(p.shape[:-1], _f)
    weight = _f(0.53333336)
    for _ in range(4):
        total = total + weight * _noise(p)
        p = p * _f(2.07) + np.array([13.1, 7.7, 3.3], _f)
        weight = weight * _f(0.5)
    return total


def _gauss(x):
    return np.exp(_f(-0.5) * x * x).astype(_f)


def color(longitude, latitude) -> np.ndarray:
    """``sky::color`` (src/sky.rs:57-101): linear RGB in [0, 1], shape (..., 3)."""
    longitude = np.asarray(longitude, _f)
    latitude = np.asarray(latitude, _f)
    longitude, latitude = np.broadcast_arrays(longitude, latitude)
    pi, tau = _f(np.pi), _f(2 * np.pi)
    lon = np.mod(longitude + pi, tau) - pi
    sin_l, cos_l = np.sin(lon), np.cos(lon)
    sin_b, cos_b = np.sin(latitude), np.cos(latitude)
    direction = np.stack([cos_l * cos_b, sin_l * cos_b, sin_b], -1).astype(_f)
    equator = np.stack([cos_l, sin_l, np.zeros_like(lon)], -1).astype(_f)
    coarse = _noise(equator * _f(3.7) + _f(17.0))
    center = _gauss(lon / _f(0.72))
    midline = _f(0.075) * np.sin(_f(2.0) * lon) + _f(0.10) * (coarse - _f(0.5))
    width = _f(0.12) + _f(0.12) * coarse + _f(0.16) * center
    warp = np.stack([_noise(direction * _f(5.0) + np.array([11.0, 0.0, 0.0], _f)),
                     _noise(direction * _f(5.0) + np.array([0.0, 19.0, 0.0], _f)),
                     _noise(direction * _f(5.0) + np.array([0.0, 0.0, 23.0], _f))], -1) - _f(0.5)
    clumps = _clouds(direction * _f(10.0) + warp * _f(4.0))
    wisps = _clouds(direction * _f(29.0) + warp * _f(7.0))
    irregular = latitude - midline + _f(0.15) * (clumps - _f(0.5))
    band = _gauss(irregular / width)
    halo = _f(0.07) * _gauss(latitude / (width * _f(1.7)))
    bulge = _f(0.46) * center * _gauss((latitude + _f(0.04)) / _f(0.28))
    starlight = band * (_f(0.05) + _f(1.35) * clumps * clumps) * (_f(0.60) + _f(0.70) * wisps) + bulge
    spine = midline + _f(0.08) * (_noise(direction * _f(8.0)) - _f(0.5))
    rift_width = _f(0.014) + _f(0.042) * _noise(direction * _f(12.0))
    rift = _gauss((latitude - spine) / rift_width)
    branch_spine = spine + _f(0.13) * np.sin(_f(3.0) * lon + _f(0.7))
    branch = _gauss((latitude - branch_spine) / _f(0.025))
    dust_clouds = _clouds(direction * _f(19.0) + warp * _f(8.0) + _f(11.0))
    patches = np.clip((_f(0.62) - dust_clouds) * _f(3.0), 0.0, 1.0)
    dust = np.clip(_f(0.82) * rift + _f(0.35) * branch * center + _f(0.65) * patches * band, 0.0, 0.96)
    warmth = np.clip(center * _f(0.85) + (coarse - _f(0.5)) * _f(0.35), 0.0, 1.0)
    cool, warm = np.array([0.38, 0.55, 1.0], _f), np.array([1.0, 0.67, 0.36], _f)
    tint = cool + (warm - cool) * warmth[..., None]
    violet = np.array([0.64, 0.32, 0.85], _f) * (halo * (_f(0.5) + clumps))[..., None]
    nebula = (np.array([0.85, 0.22, 0.38], _f)
              * (band * np.maximum((wisps - _f(0.58)) * _f(2.0), 0.0) * (_f(0.3) + center))[..., None])
    rgb = (tint * starlight[..., None] + nebula) * (_f(1.0) - dust)[..., None] + violet
    rgb = np.clip(rgb, 0.0, 1.0).astype(_f)
    return np.where((np.abs(latitude) > _f(1.35))[..., None], _f(0.0), rgb)


@functools.lru_cache(maxsize=1)
def panorama() -> np.ndarray:
    """``sky::panorama`` (src/sky.rs:103-117): (HEIGHT, WIDTH) packed RGB10."""
    y = np.arange(HEIGHT, dtype=_f)
    x = np.arange(WIDTH, dtype=_f)
    latitude = _f(np.pi) * (_f(0.5) - (y + _f(0.5)) / _f(HEIGHT))
    longitude = _f(2 * np.pi) * (_f(0.5) - (x + _f(0.5)) / _f(WIDTH))
    rgb = color(longitude[None, :], latitude[:, None])
    q = np.round(rgb * _f(1023.0)).astype(np.uint32)
    packed = q[..., 0] | (q[..., 1] << np.uint32(10)) | (q[..., 2] << np.uint32(20))
    packed.setflags(write=False)
    return packed
