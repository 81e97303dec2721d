"""Ground scenery in the reference's primitive format (src/ground.rs).

The reference builds the observer's landscape on the host: rocks, conifers,
undergrowth, logs and a cottage around the forest lookout
(``landscape_with_density``, src/ground.rs:236-433) and tumbled rocks along
the observing line for the other ground themes (src/ground.rs:83-158),
packed into its flat skip-count hierarchy (``cluster``, src/ground.rs:202-225).
These builders produce the same primitives as numpy ``PRIMITIVE`` records in
metres (east, height, north), for ``Frame.ground``.

The ground surface under each primitive comes from a ``surface_height(x, z)``
callable (metres above the tangent plane, without the curvature drop, which
is added here). The reference samples its generated terrain there
(``crate::terrain``); with ``None`` the surface is the host sphere itself.
"""
from __future__ import annotations

import math

import numpy as np

from .layouts import PRIMITIVE

CAPACITY = 256  # src/ground.rs:8
TUMBLED_ROCK = 2.25  # src/ground.rs:11

# Primitive kinds (center.w) and materials (albedo.w), src/ground.rs:167-169.
KIND_BOX, KIND_ELLIPSOID, KIND_ROCK, KIND_CONIFER, KIND_GRASS, KIND_LOG = 0.0, 1.0, 2.0, 3.0, 4.0, 5.0
MATERIAL_SOLID, MATERIAL_STONE, MATERIAL_FOLIAGE, MATERIAL_BARK, MATERIAL_GRASS = 0.0, 2.0, 3.0, 4.0, 5.0

THEME_FOREST, THEME_DESERT, THEME_CRATERS, THEME_ALPINE, THEME_BASALT, THEME_DUNES = range(6)

# Rock tints per theme (src/ground.rs:89-113); the desert's are the default.
_THEME_TINTS = {
    THEME_CRATERS: ((0.20, 0.195, 0.19), (0.13, 0.128, 0.125), (0.26, 0.25, 0.24), (0.16, 0.155, 0.15)),
    THEME_BASALT: ((0.06, 0.055, 0.052), (0.035, 0.033, 0.032), (0.10, 0.06, 0.045), (0.05, 0.045, 0.043)),
}
_DESERT_TINTS = ((0.36, 0.25, 0.15), (0.19, 0.17, 0.14), (0.43, 0.36, 0.26), (0.27, 0.13, 0.065))

F32 = np.float32


def shape(center, extent, albedo, kind: float, angle: float) -> np.ndarray:
    """One primitive with material 0 (src/ground.rs:195-201)."""
    p = np.zeros((), PRIMITIVE)
    p["center"] = [center[0], center[1], center[2], kind]
    p["extent"] = [extent[0], extent[1], extent[2], angle]
    p["albedo"] = [albedo[0], albedo[1], albedo[2], 0.0]
    return p


def material(p: np.ndarray, kind: float) -> np.ndarray:
    p["albedo"][3] = kind
    return p


def rotation(p: np.ndarray) -> np.ndarray:
    """Primitive::rotation (src/ground.rs:173-186): a turn about the north
    axis by extent.w, or for a tumbled rock the packed 8-bit Euler angles
    x, y, z applied as Rz Ry Rx."""
    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])

    if float(p["center"][3]) != TUMBLED_ROCK:
        return rz(float(p["extent"][3]))
    bits = int(float(p["extent"][3]))
    step = 2.0 * math.pi / 256.0
    x, y, z = (bits & 255) * step, ((bits >> 8) & 255) * step, ((bits >> 16) & 255) * step
    cy, sy, cx, sx = math.cos(y), math.sin(y), math.cos(x), math.sin(x)
    ry = np.array([[cy, 0.0, sy], [0.0, 1.0, 0.0], [-sy, 0.0, cy]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cx, -sx], [0.0, sx, cx]])
    return rz(z) @ ry @ rx


def world_extent(p: np.ndarray) -> np.ndarray:
    """Half extents of the rotated box (src/ground.rs:188-193)."""
    r = np.abs(rotation(p))
    return r @ p["extent"][:3].astype(np.float64)


def cluster(out: list, children: list) -> None:
    """Append a group marker bounding ``children`` (plus 0.02 m), then them
    (src/ground.rs:202-225)."""
    if not children:
        return
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for c in children:
        e = world_extent(c)
        centre = c["center"][:3].astype(np.float64)
        lo = np.minimum(lo, centre - e)
        hi = np.maximum(hi, centre + e)
    out.append(shape((lo + hi) * 0.5, (hi - lo) * 0.5 + 0.02, (0.0, 0.0, 0.0), -float(len(children)), 0.0))
    out.extend(children)


def _surface(radius_m: float, surface_height):
    def surface(x: float, z: float) -> float:
        d2 = float(x) ** 2 + float(z) ** 2
        drop = -d2 / (radius_m + math.sqrt(max(radius_m * radius_m - d2, 0.0)))
        return float(F32(drop) + F32(surface_height(x, z) if surface_height else 0.0))
    return surface


def landscape_with_density(radius_m: float, density: int = 100, surface_height=None) -> np.ndarray:
    """The forest lookout (src/ground.rs:236-433): four outcrops of stone
    (kind 2), eight groves of conifers (kind 3, foliage), four patches of
    undergrowth (kind 4, grass), two logs (kind 5, bark) and the cottage
    (boxes), each group clustered, the whole clustered once more. ``density``
    (0-100) thins trees and undergrowth as the reference does."""
    density = min(int(density), 100)
    surface = _surface(radius_m, surface_height)
    landscape: list = []
    for cx, cz, size in ((-8.0, 14.0, 0.75), (11.0, 26.0, 1.0), (-35.0, 68.0, 2.6), (53.0, 95.0, 3.2)):
        rocks = []
        for j in range(4):
            n = F32(j)
            x = F32(cx + math.sin(n * 2.4) * size * 1.6)
            z = F32(cz + math.cos(n * 3.7) * size)
            h = F32(size * (0.5 + 0.5 * abs(math.sin(n * 1.8 + 0.4))))
            rocks.append(material(shape((x, surface(x, z) + h * 0.38, z),
                                        (h * (1.1 + 0.3 * math.cos(n * 1.3)), h, h * 0.85),
                                        (0.13, 0.125, 0.115), KIND_ROCK, 0.17 * (n - 1.5)), MATERIAL_STONE))
        cluster(landscape, rocks)
    for cx, cz, count in ((-33.0, 72.0, 9), (50.0, 110.0, 10), (-110.0, 210.0, 13), (145.0, 270.0, 13),
                          (-240.0, 530.0, 15), (285.0, 590.0, 15), (-65.0, -100.0, 10), (80.0, -160.0, 10)):
        trees = []
        for i in range(count):
            if (i * 37 + 17) % 100 >= density:
                continue
            n = F32(i)
            phase = n * 2.39996 + cx * 0.07
            spread = 7.5 * math.sqrt(n + 0.5)
            x = F32(cx + spread * math.cos(phase))
            z = F32(cz + spread * math.sin(phase))
            if z > 0.0 and abs(x) < z * 0.24:
                continue  # the observing gap
            h = F32(9.0 + 6.0 * abs(math.sin(n * 7.13 + cx)))
            width = F32(h * (0.18 + 0.035 * math.cos(n * 3.4)))
            trees.append(material(shape((x, surface(x, z) + h * 0.5, z), (width, h * 0.5, width * 0.9),
                                        (0.045 + 0.012 * abs(math.sin(n * 2.1)), 0.095, 0.042), KIND_CONIFER,
                                        0.018 * math.sin(n * 1.3)), MATERIAL_FOLIAGE))
        cluster(landscape, trees)
    for cx, cz in ((-5.5, 8.0), (8.0, 18.0), (-14.0, 29.0), (21.0, 42.0)):
        plants = []
        for i in range(12):
            if (i * 37 + 17) % 100 >= density:
                continue
            n = F32(i)
            angle = n * 2.39996
            spread = 0.9 * math.sqrt(n + 0.5)
            x = F32(cx + spread * math.cos(angle))
            z = F32(cz + spread * math.sin(angle))
            h = F32(0.22 + 0.24 * abs(math.sin(n * 3.1)))
            plants.append(material(shape((x, surface(x, z) + h * 0.5, z), (0.30, h * 0.5, 0.30),
                                         (0.12, 0.18, 0.043), KIND_GRASS, 0.0), MATERIAL_GRASS))
        cluster(landscape, plants)
    wood = (0.095, 0.055, 0.026)
    for x, z, length in ((-9.0, 20.0, 2.6), (17.0, 42.0, 1.8)):
        landscape.append(material(shape((x, surface(x, z) + 0.22, z), (0.25, length, 0.25), wood, KIND_LOG, 1.43),
                                  MATERIAL_BARK))
    # The cottage east of the observing view: walls, two roof slabs meeting
    # at a ridge, gable boards, a door, two framed windows and a chimney.
    objects = []
    x, z = 22.0, 30.0
    base = surface(x, z)
    plaster, roof = (0.42, 0.34, 0.24), (0.16, 0.085, 0.055)
    objects.append(shape((x, base + 1.7, z), (3.2, 1.7, 3.8), plaster, KIND_BOX, 0.0))
    pitch = math.radians(30.0)
    for side in (-1.0, 1.0):
        objects.append(shape((x + side * 1.85, base + 4.39, z), (2.16, 0.13, 4.15), roof, KIND_BOX, -side * pitch))
    for row in range(8):
        y = 3.4 + (row + 0.5) * 0.23
        half_width = (5.46 - y) / math.tan(pitch)
        objects.append(shape((x, base + y, z), (min(half_width, 3.2), 0.12, 3.8), plaster, KIND_BOX, 0.0))
    objects.append(shape((x - 0.5, base + 1.05, z - 3.82), (0.55, 1.05, 0.035), wood, KIND_BOX, 0.0))
    for dx in (-2.0, 1.6):
        objects.append(shape((x + dx, base + 2.0, z - 3.83), (0.55, 0.55, 0.04), wood, KIND_BOX, 0.0))
        objects.append(shape((x + dx, base + 2.0, z - 3.88), (0.45, 0.45, 0.015), (0.10, 0.19, 0.23), KIND_BOX, 0.0))
        objects.append(shape((x + dx, base + 2.0, z - 3.90), (0.035, 0.46, 0.015), wood, KIND_BOX, 0.0))
    objects.append(shape((x + 1.5, base + 5.0, z + 1.8), (0.35, 1.2, 0.45), (0.32, 0.20, 0.14), KIND_BOX, 0.0))
    cluster(landscape, objects)
    bounded: list = []
    cluster(bounded, landscape)
    assert len(bounded) <= CAPACITY
    return np.array(bounded, PRIMITIVE)


def theme_rocks(radius_m: float, theme: int, surface_height=None) -> np.ndarray:
    """Scattered tumbled rocks along the observing line for the desert,
    crater and basalt grounds (src/ground.rs:83-158): 28 rocks alternating
    sides of the line, four proportions (slabs, fragments, squat stones,
    boulders), packed Euler angles, partly buried, clustered once."""
    if theme in (THEME_FOREST, THEME_ALPINE, THEME_DUNES):
        raise ValueError("the forest has landscape_with_density; the alpine and dune grounds carry no rocks")
    tints = _THEME_TINTS.get(theme, _DESERT_TINTS)
    surface = _surface(radius_m, surface_height)
    seed = 0xA7ACA4A1

    def random() -> float:
        nonlocal seed
        seed ^= (seed << 13) & 0xFFFFFFFF
        seed ^= seed >> 17
        seed ^= (seed << 5) & 0xFFFFFFFF
        return float(F32((seed >> 8) / 16_777_216.0))

    proportions = ((1.8, 0.45, 1.1), (0.7, 0.75, 1.9), (1.05, 0.8, 0.9), (0.9, 1.25, 1.15))
    rocks = []
    for i in range(28):
        n = float(i)
        side = -1.0 if i % 2 == 0 else 1.0
        z = F32(-9.0 + n * 6.5 + (random() - 0.5) * 5.0)
        x = F32(side * (9.0 + abs(math.sin(n * 2.4)) * 24.0))
        h = 0.09 + 0.70 * random() ** 2
        extent = [v * h * (0.75 + random() * 0.5) for v in proportions[i % 4]]
        angles = [int(random() * 256.0) for _ in range(3)]
        orientation = float(angles[0] | (angles[1] << 8) | (angles[2] << 16))
        tint = tints[i % 4]
        shade = 0.75 + random() * 0.45
        rock = material(shape((x, surface(x, z), z), extent, [v * shade for v in tint], TUMBLED_ROCK, orientation),
                        MATERIAL_STONE)
        # Embed the bottom in the ground by part of the rotated vertical extent.
        rock["center"][1] += world_extent(rock)[1] * (0.12 + random() * 0.24)
        rocks.append(rock)
    out: list = []
    cluster(out, rocks)
    return np.array(out, PRIMITIVE)
