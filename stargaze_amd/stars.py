"""Catalogue stars in the tracer's ray format.

Restates the host side of stargaze's background star field: the radiance of
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
