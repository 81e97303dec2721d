"""Heightfield data for the terrain quadtree, in the reference's buffer layout.

``Terrain::data`` (src/terrain.rs:47-52, built at :781-837) is
``[heights W*W | x-derivatives W*W | z-derivatives W*W | min,max per node]``
with the derivative blocks present only for grids of at most 1024 cells, and
node bounds in heap order (node k's children are 4k+1..4k+4, child c offset
by (c & 1, c >> 1) half-tiles). The HIP library turns this into cached,
curved AABBs and vertices on the GPU (cache_terrain_geometry).

Two sources of heights: ``generate`` restates the reference's own generator
for each ground theme (``Terrain::generate``, src/terrain.rs:83-790: warped
multifractal noise with talus erosion, craters, star dunes, and the surveyed
ice field and volcanoes), and ``ridged_fbm`` is a seeded generator of this
build for the synthetic benchmark workloads.
"""
from __future__ import annotations

import numpy as np

EXTENT = 16_000.0  # src/terrain.rs:13


def coordinates(cells: int) -> np.ndarray:
    """``coordinate_in`` (src/terrain.rs:75-78): quadratic vertex spacing, metres."""
    i = np.arange(cells + 1, dtype=np.float32)
    u = np.float32(2.0) * i / np.float32(cells) - np.float32(1.0)
    return (u * np.abs(u) * np.float32(EXTENT)).astype(np.float32)


def scale(radius_m: float) -> float:
    """``terrain::scale`` (src/terrain.rs:80-82)."""
    return float(np.clip(radius_m / 200_000.0, 0.0001, 1.0))


def _lattice(ix, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B1) ^ iz.astype(np.uint64) * np.uint64(0x85EBCA77)
         ^ np.uint64(seed)) & np.uint64(0xFFFFFFFF)
    h = h.astype(np.uint32)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h.astype(np.float32) / np.float32(4294967296.0)


def _value_noise(x, z, seed):
    ix, iz = np.floor(x), np.floor(z)
    fx, fz = x - ix, z - iz
    ux, uz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    ix, iz = ix.astype(np.int64), iz.astype(np.int64)
    a = _lattice(ix, iz, seed)
    b = _lattice(ix + 1, iz, seed)
    c = _lattice(ix, iz + 1, seed)
    d = _lattice(ix + 1, iz + 1, seed)
    return (a * (1 - ux) + b * ux) * (1 - uz) + (c * (1 - ux) + d * ux) * uz


def ridged_fbm(cells: int, seed: int = 0x743337A9, octaves: int = 8, amplitude: float = 800.0,
               wavelength: float = 6000.0) -> np.ndarray:
    """Ridged fBm heights (metres) on the quadratic grid, shape (W, W) [z, x].

    The observer's lookout at the grid centre is levelled to height 0, as on
    the reference's forest ground.
    """
    coord = coordinates(cells).astype(np.float64)
    x = coord[None, :]
    z = coord[:, None]
    total = np.zeros((cells + 1, cells + 1), np.float64)
    amp, freq, norm = 1.0, 1.0 / wavelength, 0.0
    weight = np.ones_like(total)
    for o in range(octaves):
        n = _value_noise(x * freq + 31.7 * o, z * freq - 17.3 * o, (seed + 0x632BE5AB * o) & 0xFFFFFFFF)
        ridge = 1.0 - np.abs(2.0 * n - 1.0)
        ridge = ridge * ridge * weight
        weight = np.clip(ridge * 2.0, 0.0, 1.0)
        total += amp * ridge
        norm += amp
        amp *= 0.5
        freq *= 2.0
    h = total / norm * amplitude
    # Level the observatory: blend to the centre height within ~60 m.
    r = np.sqrt(x * x + z * z)
    centre = h[cells // 2, cells // 2]
    blend = np.clip((r - 20.0) / 40.0, 0.0, 1.0)
    h = (h - centre) * blend
    return h.astype(np.float32)


def _spread(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint32) & np.uint32(0x0000FFFF)
    v = (v | (v << np.uint32(8))) & np.uint32(0x00FF00FF)
    v = (v | (v << np.uint32(4))) & np.uint32(0x0F0F0F0F)
    v = (v | (v << np.uint32(2))) & np.uint32(0x33333333)
    v = (v | (v << np.uint32(1))) & np.uint32(0x55555555)
    return v


def node_count(cells: int) -> int:
    leaves = cells // 2
    return (4 * leaves * leaves - 1) // 3


def bounds_offset(cells: int) -> int:
    return (cells + 1) * (cells + 1) * (1 if cells > 1024 else 3)


def terrain_data(heights: np.ndarray, cells: int) -> np.ndarray:
    """``Terrain::data`` for ``heights`` (src/terrain.rs:781-837)."""
    w = cells + 1
    assert heights.shape == (w, w) and cells >= 4 and cells & (cells - 1) == 0
    h = heights.astype(np.float32)
    out = np.zeros(bounds_offset(cells) + 2 * node_count(cells), np.float32)
    out[: w * w] = h.reshape(-1)
    if cells <= 1024:
        coord = coordinates(cells)
        idx = np.arange(w)
        lo = np.maximum(idx - 1, 0)
        hi = np.minimum(idx + 1, cells)
        dx = (h[:, hi] - h[:, lo]) / (coord[hi] - coord[lo])[None, :]
        dz = (h[hi, :] - h[lo, :]) / (coord[hi] - coord[lo])[:, None]
        out[w * w: 2 * w * w] = dx.reshape(-1)
        out[2 * w * w: 3 * w * w] = dz.reshape(-1)
    # Leaves: min/max over the 3x3 vertices of each 2x2-cell tile.
    leaves = cells // 2
    view = np.lib.stride_tricks.sliding_window_view(h, (3, 3))[::2, ::2]
    lo_lvl = view.min(axis=(2, 3))
    hi_lvl = view.max(axis=(2, 3))
    assert lo_lvl.shape == (leaves, leaves)
    offset = bounds_offset(cells)
    level = int(np.log2(leaves))
    while True:
        n = lo_lvl.shape[0]
        tz, tx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        morton = _spread(tx) | (_spread(tz) << np.uint32(1))
        index = ((4 ** level - 1) // 3 + morton.astype(np.int64)).reshape(-1)
        out[offset + 2 * index] = lo_lvl.reshape(-1)
        out[offset + 2 * index + 1] = hi_lvl.reshape(-1)
        if n == 1:
            break
        lo_lvl = lo_lvl.reshape(n // 2, 2, n // 2, 2).min(axis=(1, 3))
        hi_lvl = hi_lvl.reshape(n // 2, 2, n // 2, 2).max(axis=(1, 3))
        level -= 1
    return out


def height(heights: np.ndarray, cells: int, x: float, z: float) -> float:
    """``Terrain::height`` (src/terrain.rs:840-866): the GPU triangles' height."""
    coord = coordinates(cells)

    def index(v):
        u = np.sign(v) * np.sqrt(abs(v) / EXTENT)
        return min(int(np.floor((u + 1.0) * 0.5 * cells)), cells - 1)

    ix, iz = index(x), index(z)
    u = float(np.clip((x - coord[ix]) / (coord[ix + 1] - coord[ix]), 0.0, 1.0))
    v = float(np.clip((z - coord[iz]) / (coord[iz + 1] - coord[iz]), 0.0, 1.0))
    a, b = heights[iz, ix], heights[iz, ix + 1]
    c, d = heights[iz + 1, ix], heights[iz + 1, ix + 1]
    if u + v <= 1.0:
        return float(a + (b - a) * u + (c - a) * v)
    return float(d + (c - d) * (1.0 - u) + (b - d) * (1.0 - v))


def heights_at(heights: np.ndarray, cells: int, x: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Vectorised ``height`` over arrays of surface positions (metres)."""
    coord = coordinates(cells).astype(np.float64)
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)

    def index(v):
        u = np.sign(v) * np.sqrt(np.abs(v) / EXTENT)
        return np.minimum(np.floor((u + 1.0) * 0.5 * cells).astype(np.int64), cells - 1)

    ix, iz = index(x), index(z)
    u = np.clip((x - coord[ix]) / (coord[ix + 1] - coord[ix]), 0.0, 1.0)
    v = np.clip((z - coord[iz]) / (coord[iz + 1] - coord[iz]), 0.0, 1.0)
    h = heights.astype(np.float64)
    a, b = h[iz, ix], h[iz, ix + 1]
    c, d = h[iz + 1, ix], h[iz + 1, ix + 1]
    lower = a + (b - a) * u + (c - a) * v
    upper = d + (c - d) * (1.0 - u) + (b - d) * (1.0 - v)
    return np.where(u + v <= 1.0, lower, upper)


# -- the reference's generator (src/terrain.rs:83-837) ------------------------
# Vectorised over the whole grid in float32, operation for operation; the
# transcendentals (sin, cos, pow, log, hypot) come from numpy's float32
# routines, so results may differ from the reference's in the last bits.
F = np.float32
SEED = 0x743337A9  # src/terrain.rs:45
LEAF = 2
THEME_CELLS = {0: 512, 1: 512, 2: 1024, 3: 1024, 4: 2048, 5: 1024}  # terrain::cells (:27-36)
CRATER_CELLS, ALPINE_CELLS, DUNE_CELLS, MAX_CELLS = 1024, 1024, 1024, 2048
ICE_FLOW = F(4.4)  # :308, mirrored by ground.wgsl
LOOKOUT_HILL = (F(180.0), F(700.0))  # :313
DUNE_PAN_LEVEL = F(2.0)  # :436
DUNE_PANS = ((F(842.6), F(707.1), F(560.0)), (F(4_346.7), F(1_164.7), F(900.0)))  # :437-442
DUNE_WIND = F(0.61)  # :622
TAU = F(2.0 * np.pi)
# Island, bearing (deg), distance (m), turn (deg), size of the volcanoes on
# the basalt ground's horizon (src/volcano.rs:5-14).
VOLCANO_SITES = ((1, 8.0, 12_300.0, 0.0, 1.2), (2, 52.0, 12_800.0, -40.0, 1.0), (0, 95.0, 11_800.0, 160.0, 0.85),
                 (1, 138.0, 13_200.0, 90.0, 0.8), (2, 182.0, 12_000.0, 30.0, 1.15), (0, 228.0, 13_000.0, 300.0, 0.9),
                 (1, 270.0, 11_600.0, 200.0, 1.05), (2, 318.0, 12_600.0, 120.0, 0.95))


def _u32(v) -> np.ndarray:
    return np.asarray(v).astype(np.uint32)


def _hash(v) -> np.ndarray:
    """lowbias32 (src/terrain.rs:83-87), wrapping u32 products."""
    v = np.atleast_1d(_u32(v))
    with np.errstate(over="ignore"):
        v = (v ^ (v >> np.uint32(16))) * np.uint32(0x7FEB352D)
        v = (v ^ (v >> np.uint32(15))) * np.uint32(0x846CA68B)
    return v ^ (v >> np.uint32(16))


def _cell_key(ix, iz, seed) -> np.ndarray:
    with np.errstate(over="ignore"):
        key = (np.atleast_1d(_u32(ix)) * np.uint32(0x8DA6B343)) ^ (np.atleast_1d(_u32(iz)) * np.uint32(0xD8163841))
    return _hash(key ^ np.uint32(seed))


def noise(x, z, seed: int) -> np.ndarray:
    """Gradient noise with eight directions, scaled by 1.5 (:89-121)."""
    x, z = F(x) if np.isscalar(x) else x, F(z) if np.isscalar(z) else z
    x, z = np.asarray(x, F), np.asarray(z, F)
    fx, fz = np.floor(x), np.floor(z)
    ix, iz = fx.astype(np.int32), fz.astype(np.int32)
    x, z = x - fx, z - fz

    def fade(v):
        return v * v * v * (v * (v * F(6.0) - F(15.0)) + F(10.0))

    r = F(0.70710677)

    def grad(dx, dz):
        h = _cell_key(ix + dx, iz + dz, seed) & np.uint32(7)
        u, v = x - F(dx), z - F(dz)
        return np.select([h == 0, h == 1, h == 2, h == 3, h == 4, h == 5, h == 6],
                         [u, -u, v, -v, (u + v) * r, (u - v) * r, (-u + v) * r], (-u - v) * r).astype(F)

    a, b, c, d = grad(0, 0), grad(1, 0), grad(0, 1), grad(1, 1)
    low = a + (b - a) * fade(x)
    high = c + (d - c) * fade(x)
    return (low + (high - low) * fade(z)) * F(1.5)


def fbm(x, z, seed: int, octaves: int) -> np.ndarray:
    """:123-135."""
    x, z = np.asarray(x, F), np.asarray(z, F)
    total, weight = np.zeros(np.broadcast(x, z).shape, F), F(0.55)
    for octave in range(octaves):
        total = total + weight * noise(x, z, (seed + octave * 71) & 0xFFFFFFFF)
        x, z = ((F(0.8) * x + F(0.6) * z) * F(2.03) + F(17.2), (F(-0.6) * x + F(0.8) * z) * F(2.03) + F(9.1))
        weight = weight * F(0.48)
    return total


def ridges(x, z, seed: int) -> np.ndarray:
    """Ridged multifractal with signal feedback (:137-153)."""
    x, z = np.asarray(x, F), np.asarray(z, F)
    total = np.zeros(np.broadcast(x, z).shape, F)
    amplitude, feedback = F(0.55), F(1.0)
    for octave in range(7):
        ridge = np.maximum(F(1.0) - np.abs(noise(x, z, (seed + octave * 109) & 0xFFFFFFFF)), F(0.0))
        signal = ridge * ridge * feedback
        total = total + amplitude * signal
        feedback = np.clip(signal * F(1.9), F(0.0), F(1.0))
        x, z = ((F(0.8) * x + F(0.6) * z) * F(2.05) + F(5.7), (F(-0.6) * x + F(0.8) * z) * F(2.05) + F(13.4))
        amplitude = amplitude * F(0.49)
    return np.clip((total - F(0.28)) * F(1.45), F(0.0), F(1.0))


def smooth(low, high, x) -> np.ndarray:
    """Smoothstep (:155-158)."""
    t = np.clip((np.asarray(x, F) - F(low)) / (F(high) - F(low)), F(0.0), F(1.0))
    return t * t * (F(3.0) - F(2.0) * t)


def _reach(x, z):
    return np.maximum(np.abs(x), np.abs(z))


def edge(x, z) -> np.ndarray:
    """Fades every surface to zero at the tile edge (:191-194)."""
    return F(1.0) - smooth(F(EXTENT) - F(2000.0), F(EXTENT), _reach(x, z))


def cell_size(x, z, cells: int) -> np.ndarray:
    """:465-470."""
    return F(4.0) * F(EXTENT) / F(cells) * np.sqrt(_reach(x, z) / F(EXTENT))


def rounding_width(x, z, cells: int) -> np.ndarray:
    return np.maximum(F(1.5) * cell_size(x, z, cells), F(1.0))


def detail(x, z, feature: float, cells: int) -> np.ndarray:
    """:473-478."""
    return F(1.0) - smooth(F(0.25) * F(feature), F(0.5) * F(feature), cell_size(x, z, cells))


def rounded(distance, x, z) -> np.ndarray:
    """:481-485."""
    width = np.maximum(F(2.4) * cell_size(x, z, DUNE_CELLS), F(1.0))
    return np.hypot(np.asarray(distance, F), width) - width


def smooth_min(a, b, k) -> np.ndarray:
    """:488-491."""
    h = np.maximum(k - np.abs(a - b), F(0.0)) / k
    return np.minimum(a, b) - h * h * k * F(0.25)


def _powi(u, n: int):
    if n == 4:
        u2 = u * u
        return u2 * u2
    if n == 6:
        u2 = u * u
        return u2 * u2 * u2
    if n == -3:
        return F(1.0) / (u * u * u)
    raise ValueError(n)


def elevation(x, z, seed: int = SEED) -> np.ndarray:
    """The forest ground: a meandering valley through warped ridges (:160-177)."""
    wx = F(1_150.0) * fbm(x / F(4_700.0) + F(5.0), z / F(4_700.0), seed, 3)
    wz = F(1_150.0) * fbm(x / F(4_700.0), z / F(4_700.0) - F(11.0), seed ^ 0x59AB, 3)
    continental = np.clip(F(0.65) + F(0.5) * fbm(x / F(8_000.0) + F(7.3), z / F(8_000.0) - F(3.9), seed, 3),
                          F(0.2), F(1.0))
    ridge = ridges((x + wx) / F(2_200.0) + F(4.2), (z + wz) / F(2_200.0) + F(1.7), seed)
    valley_center = F(850.0) * np.sin(z / F(2_800.0)) + F(230.0) * np.sin(z / F(950.0))
    valley = smooth(80.0, 2_000.0, np.abs(x - valley_center))
    mountain = F(900.0) * continental * np.power(ridge, F(1.35)) * valley
    floor = (F(9.0) + F(20.0) * fbm(x / F(430.0), z / F(430.0), seed ^ 0x3847, 5)
             + F(2.0) * fbm(x / F(34.0), z / F(34.0), seed, 3))
    clearing = smooth(12.0, 85.0, np.hypot(x, z))
    fade = F(1.0) - smooth(F(EXTENT) - F(2_000.0), F(EXTENT), _reach(x, z))
    return np.maximum(mountain + floor, F(0.0)) * clearing * fade


def desert_elevation(x, z, seed: int = SEED) -> np.ndarray:
    """A barren alluvial basin between broken ridges (:179-189)."""
    basin = smooth(450.0, 2400.0, np.abs(x + F(150.0) * np.sin(z / F(3000.0))))
    hills = F(700.0) * basin * ridges(x / F(2800.0) + F(2.1), z / F(3300.0) + F(7.0), seed)
    plain = (F(3.0) + F(5.0) * fbm(x / F(650.0), z / F(650.0), seed, 4)
             + F(0.35) * fbm(x / F(25.0), z / F(25.0), seed ^ 0x1234, 3))
    fade = F(1.0) - smooth(F(EXTENT) - F(2000.0), F(EXTENT), _reach(x, z))
    return np.maximum(hills + plain, F(0.0)) * smooth(15.0, 90.0, np.hypot(x, z)) * fade


def craters(x, z, cell: float, seed: int) -> np.ndarray:
    """Grid-jittered craters of one size class: bowl, rim and ejecta (:197-232)."""
    cell = F(cell)
    cx, cz = np.floor(x / cell).astype(np.int32), np.floor(z / cell).astype(np.int32)
    total = np.zeros(np.broadcast(x, z).shape, F)
    width = F(1.2) * rounding_width(x, z, CRATER_CELLS)
    for dz in (-1, 0, 1):
        for dx in (-1, 0, 1):
            ix, iz = cx + dx, cz + dz
            h = _cell_key(ix, iz, seed)

            def unit(shift):
                return ((h >> np.uint32(shift)) & np.uint32(1023)).astype(F) / F(1023.0)

            keep = unit(30) <= F(0.75)
            centre_x = (ix.astype(F) + F(0.15) + F(0.7) * unit(0)) * cell
            centre_z = (iz.astype(F) + F(0.15) + F(0.7) * unit(10)) * cell
            radius = cell * (F(0.12) + F(0.3) * unit(20))
            distance = np.hypot(x - centre_x, z - centre_z)
            u = distance / radius
            depth, rim = F(0.2) * radius, F(0.1) * radius
            reach = F(1.0) - smooth(F(0.6) * cell, F(1.1) * cell, distance)
            inner = depth * (u * u - F(1.0)) + rim * _powi(u, 4)
            outer = rim * _powi(np.maximum(u, F(1e-3)), -3) * reach
            total = total + np.where(keep, smooth_min(inner, outer, width), F(0.0))
    return total


def crater_elevation(x, z, seed: int = SEED) -> np.ndarray:
    """A regolith plain with craters and the lookout on a fresh rim (:236-261)."""
    plain = (F(60.0) + F(10.0) * fbm(x / F(900.0), z / F(900.0), seed, 4)
             + F(0.6) * fbm(x / F(30.0), z / F(30.0), seed ^ 0x77, 3) * detail(x, z, 30.0, CRATER_CELLS))
    pits = (craters(x, z, 2600.0, seed ^ 0x1001) + craters(x, z, 700.0, seed ^ 0x2002)
            + F(0.6) * craters(x, z, 160.0, seed ^ 0x3003))
    u = np.hypot(x, z - F(950.0)) / F(950.0)
    landmark = smooth_min(F(170.0) * (u * u - F(1.0)) + F(75.0) * _powi(u, 6),
                          F(75.0) * _powi(np.maximum(u, F(1e-3)), -3), F(1.2) * rounding_width(x, z, CRATER_CELLS))
    lookout = smooth(15.0, 90.0, np.hypot(x, z))
    return np.maximum(plain + landmark + pits * lookout, F(0.0)) * edge(x, z)


def dem_sample(data: np.ndarray, base: int, samples: int, fx, fz) -> np.ndarray:
    """Catmull-Rom interpolation of a square grid of decimetre heights
    (:263-291); ``data`` is the little-endian u16 file as an array."""
    i, j = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)

    def weights(t):
        t2, t3 = t * t, t * t * t
        return (F(-0.5) * t3 + t2 - F(0.5) * t, F(1.5) * t3 - F(2.5) * t2 + F(1.0),
                F(-1.5) * t3 + F(2.0) * t2 + F(0.5) * t, F(0.5) * t3 - F(0.5) * t2)

    wx, wz = weights(fx - i.astype(F)), weights(fz - j.astype(F))
    total = np.zeros(np.shape(fx), F)
    for b in range(4):
        jj = np.clip(j + b - 1, 0, samples - 1)
        for a in range(4):
            ii = np.clip(i + a - 1, 0, samples - 1)
            height = data[base + jj * samples + ii].astype(F) * F(0.1)
            total = total + wx[a] * wz[b] * height
    return np.maximum(total, F(0.0))


class Surveys:
    """The reference's elevation surveys (``assets/terrain/*.u16``), read
    from a folder the caller names: the Patagonian ice field (641^2, 50 m),
    six Aeolian islands (6 x 320^2, 25 m) and Stromboli (1281^2 at 5 m, then
    a 1025^2 summit grid at 1 m)."""

    def __init__(self, folder):
        from pathlib import Path

        folder = Path(folder)
        self.patagonia = np.fromfile(folder / "patagonia-ice-field.u16", "<u2")
        self.aeolian = np.fromfile(folder / "aeolian-islands.u16", "<u2")
        self.stromboli = np.fromfile(folder / "stromboli-2023.u16", "<u2")
        assert self.patagonia.size == 641 * 641 and self.aeolian.size == 6 * 320 * 320
        assert self.stromboli.size == 1281 * 1281 + 1025 * 1025


def lookout_hill(r) -> np.ndarray:
    """The nunatak under the alpine lookout (:319-322)."""
    u = np.minimum(np.hypot(r, F(30.0)) / LOOKOUT_HILL[1], F(1.0))
    return LOOKOUT_HILL[0] * (F(1.0) - u) * (F(1.0) - u)


def _patagonia(surveys: Surveys, x, z):
    half = F(320.0)
    fx = np.clip(x / F(50.0) + half, F(0.0), F(2.0) * half)
    fz = np.clip(z / F(50.0) + half, F(0.0), F(2.0) * half)
    return dem_sample(surveys.patagonia, 0, 641, fx, fz)


def alpine_elevation(x, z, surveys: Surveys, seed: int = SEED) -> np.ndarray:
    """The Patagonian ice field around a nunatak (:324-358)."""
    cells = ALPINE_CELLS
    relative = np.maximum(_patagonia(surveys, x, z) - _patagonia(surveys, np.zeros(1, F), np.zeros(1, F))[0],
                          F(0.0))
    r = np.hypot(x, z)
    hill = lookout_hill(r)
    rock = smooth(20.0, 200.0, relative)
    crags = F(70.0) * np.power(ridges(x / F(380.0) + F(1.3), z / F(380.0) - F(2.1), seed ^ 0xA1), F(1.5)) - F(18.0)
    broken = F(7.0) * fbm(x / F(70.0), z / F(70.0), seed ^ 0xA2, 3)
    ribs = (smooth(10.0, 80.0, hill) * (F(16.0) * ridges(x / F(140.0) + F(7.0), z / F(140.0), seed ^ 0xA5) - F(13.0))
            * detail(x, z, 60.0, cells))
    mountains = relative + ribs + rock * (crags * detail(x, z, 180.0, cells) + broken * detail(x, z, 70.0, cells))
    fx, fz = np.sin(ICE_FLOW), np.cos(ICE_FLOW)
    along, across = x * fx + z * fz, x * fz - z * fx
    swell = (F(5.0) + F(4.0) * fbm(x / F(1_800.0), z / F(1_800.0), seed ^ 0xA3, 3)
             + F(1.2) * fbm(along / F(260.0), across / F(70.0), seed ^ 0xA4, 2) * detail(x, z, 70.0, cells))
    fade = F(1.0) - smooth(F(EXTENT) - F(1_000.0), F(EXTENT), _reach(x, z))
    return (np.maximum(mountains + swell, F(0.0)) * smooth(12.0, 85.0, r) + hill) * fade


def _aeolian(surveys: Surveys, index: int, x, z):
    half = F(159.5)
    reach = half * F(25.0)
    inside = (np.abs(x) < reach) & (np.abs(z) < reach)
    base = index * 320 * 320
    value = (dem_sample(surveys.aeolian, base, 320, x / F(25.0) + half, z / F(25.0) + half)
             * (F(1.0) - smooth(reach - F(300.0), reach, _reach(x, z))))
    return np.where(inside, value, F(0.0))


def _stromboli(surveys: Surveys, x, z):
    reach = _reach(x, z)
    coarse = dem_sample(surveys.stromboli, 0, 1281, x / F(5.0) + F(640.0), z / F(5.0) + F(640.0))
    fine_z = z - F(250.0)
    fine_reach = _reach(x, fine_z)
    fine = dem_sample(surveys.stromboli, 1281 * 1281, 1025, x + F(512.0), fine_z + F(512.0))
    blended = coarse + (fine - coarse) * (F(1.0) - smooth(440.0, 512.0, fine_reach))
    return np.where(reach >= F(3200.0), F(0.0), np.where(fine_reach >= F(512.0), coarse, blended))


def basalt_elevation(x, z, surveys: Surveys, seed: int = SEED) -> np.ndarray:
    """Stromboli's surveyed summit over a lava plain, with surveyed islands
    on the horizon (:416-452)."""
    volcanoes = _stromboli(surveys, x, z)
    rad = F(np.pi) / F(180.0)
    for index, bearing, distance, turn, size in VOLCANO_SITES:
        b, t = F(bearing) * rad, F(turn) * rad
        dx = (x - F(distance) * np.sin(b)) / F(size)
        dz = (z - F(distance) * np.cos(b)) / F(size)
        c, s = np.cos(t), np.sin(t)
        lx, lz = c * dx - s * dz, s * dx + c * dz
        real = _stromboli(surveys, lx, lz) if index == 0 else _aeolian(surveys, index, lx, lz)
        volcanoes = np.maximum(volcanoes, F(size) * real)
    rough = (F(4.0) * fbm(x / F(600.0), z / F(600.0), seed, 3)
             + F(2.5) * np.abs(fbm(x / F(45.0), z / F(45.0), seed ^ 0xBEEF, 3)) * detail(x, z, 45.0, MAX_CELLS)
             + F(0.8) * fbm(x / F(9.0), z / F(9.0), seed ^ 0x0F, 2) * detail(x, z, 9.0, MAX_CELLS))
    sea = F(20.0) + rough * (F(1.0) - smooth(20.0, 100.0, volcanoes))
    fade = F(1.0) - smooth(F(EXTENT) - F(1_000.0), F(EXTENT), _reach(x, z))
    return np.maximum(sea + volcanoes, F(0.0)) * fade


def dune_arm(x, z, start, bearing, length, height, slopes, phase) -> np.ndarray:
    """A sinuous sand ridge from ``start`` along ``bearing`` (:498-521)."""
    dir_x, dir_z = np.sin(bearing), np.cos(bearing)
    dx, dz = x - start[0], z - start[1]
    along = dx * dir_x + dz * dir_z
    t = np.clip(along / length, F(0.0), F(1.0))
    wander = F(140.0) * np.sin(t * F(5.0) + phase) * t
    across = dx * dir_z - dz * dir_x - wander
    crest = (height * np.power(F(1.0) - t, F(0.8)) - F(0.55) * np.maximum(along - length, F(0.0))
             - F(0.55) * np.maximum(-along, F(0.0)))
    slope = np.where(across < F(0.0), F(slopes[0]), F(slopes[1]))
    value = np.maximum(crest - slope * rounded(across, x, z), F(0.0))
    return np.where((along < F(-400.0)) | (along > length + F(400.0)), F(0.0), value)


def star_dune(x, z, cx, cz, height, seed) -> np.ndarray:
    """Three to five arms radiating from a high centre (:524-545); ``seed``
    may vary per point."""
    def unit(shift):
        return (_hash(_u32(seed) + np.uint32(shift)) >> np.uint32(8)).astype(F) / F(16_777_216.0)

    arms = 3 + (unit(1) * F(2.99)).astype(np.int32)
    spin = unit(2) * TAU
    h = np.maximum(height - F(0.5) * rounded(np.hypot(x - cx, z - cz), x, z), F(0.0))
    for k in range(5):
        bearing = spin + F(k) * TAU / arms.astype(F) + (unit(10 + k) - F(0.5)) * F(0.9)
        length = F(700.0) + F(1_100.0) * unit(20 + k)
        arm = dune_arm(x, z, (cx, cz), bearing, length, height, (0.62, 0.3), unit(30 + k) * F(6.0))
        h = np.where(k < arms, np.maximum(h, arm), h)
    return h


def dune_elevation(x, z, seed: int = SEED) -> np.ndarray:
    """Big Daddy above Deadvlei among star dunes (:547-616)."""
    rad = F(np.pi) / F(180.0)
    h = np.maximum(F(325.0) - F(0.5) * rounded(np.hypot(x, z), x, z), F(0.0))
    for bearing, length, height, phase in ((320.0, 3_200.0, 325.0, 0.0), (140.0, 2_600.0, 325.0, 2.0),
                                           (225.0, 2_200.0, 290.0, 4.0)):
        b = F(bearing) * rad
        slopes = (0.3, 0.62) if np.sin(b) < 0.0 else (0.62, 0.3)
        h = np.maximum(h, dune_arm(x, z, (F(0.0), F(0.0)), b, F(length), F(height), slopes, F(phase)))
    cell = F(1_900.0)
    gx, gz = np.floor(x / cell).astype(np.int32), np.floor(z / cell).astype(np.int32)
    for dz in (-1, 0, 1):
        for dx in (-1, 0, 1):
            ix, iz = gx + dx, gz + dz
            own = (((ix == 0) & (iz == 0)) | ((ix == -1) & (iz == -1)) | ((ix == -1) & (iz == 0))
                   | ((ix == 0) & (iz == -1)))
            key = _cell_key(ix, iz, seed)

            def unit(shift):
                return ((key >> np.uint32(shift)) & np.uint32(255)).astype(F) / F(255.0)

            cx = (ix.astype(F) + F(0.2) + F(0.6) * unit(0)) * cell
            cz = (iz.astype(F) + F(0.2) + F(0.6) * unit(8)) * cell
            star = star_dune(x, z, cx, cz, F(110.0) + F(160.0) * unit(16), key)
            h = np.where(own, h, np.maximum(h, star))
    warp = (F(450.0) * fbm(x / F(1_600.0), z / F(1_600.0) + F(9.0), seed, 3)
            + F(120.0) * fbm(x / F(400.0), z / F(400.0) - F(3.0), seed ^ 0x15, 2))
    wave = F(0.5) + F(0.5) * np.cos((x * F(0.35) + z * F(0.94) + warp) / F(680.0) * TAU)
    strength = np.clip(F(0.5) + F(0.5) * fbm(x / F(2_500.0), z / F(2_500.0), seed ^ 0x3C, 2), F(0.0), F(1.0))
    ridge = F(40.0) * wave * wave * strength
    floor = (F(7.0) + F(3.0) * fbm(x / F(900.0), z / F(900.0), seed, 3)
             + F(0.5) * fbm(x / F(40.0), z / F(40.0), seed ^ 0x6D, 3) * detail(x, z, 40.0, DUNE_CELLS))
    height = np.maximum(h, floor + np.maximum(ridge, F(0.0)))
    for px, pz, radius in DUNE_PANS:
        shore = radius * (F(1.0) + F(0.3) * fbm(x / F(350.0) + px, z / F(350.0) + pz, seed ^ 0x77, 3))
        flat = smooth(shore, shore + F(140.0), np.hypot(x - px, z - pz))
        height = height * flat + DUNE_PAN_LEVEL * (F(1.0) - flat)
    return np.maximum(height, F(0.0)) * edge(x, z)


def small_dunes(x, z, wavelength: float, height, seed: int) -> np.ndarray:
    """Transverse dunes on the large ones (:628-668)."""
    w = F(wavelength)
    wx, wz = np.sin(DUNE_WIND), np.cos(DUNE_WIND)
    along, across = x * wx + z * wz, x * wz - z * wx
    warp = (F(0.9) * w * fbm(x / (F(4.0) * w), z / (F(4.0) * w), seed, 3)
            + F(0.25) * w * np.abs(np.sin(across / (F(1.3) * w)
                                          + F(2.0) * noise(x / F(3_000.0), z / F(3_000.0), seed ^ 5))))
    phase = (along - warp) / w
    f = phase - np.floor(phase)
    lee = F(0.82)
    distance = np.where(f < lee, (lee - f) * w, (f - lee) * w)
    slope = np.where(f < lee, height / (lee * w), height / ((F(1.0) - lee) * w))
    crest = np.maximum(height - slope * rounded(distance, x, z), F(0.0))
    segments = smooth(-0.15, 0.45, fbm(across / (F(2.2) * w) + F(3.0) * np.floor(phase), along / (F(6.0) * w),
                                        seed ^ 0x2B, 2) + F(0.15))
    return crest * segments * detail(x, z, F(0.5) * w, DUNE_CELLS)


def superimpose_dunes(heights: np.ndarray, coords: np.ndarray, seed: int = SEED) -> np.ndarray:
    """Small dunes on gentle ground only (:670-709); returns new heights."""
    n = coords.size
    base = heights
    idx = np.arange(n)
    left, right = np.maximum(idx - 1, 0), np.minimum(idx + 1, n - 1)

    def steps(values, lo, hi, axis):
        gap = np.maximum(np.abs(coords[idx] - coords[lo]), F(1e-3))
        gap_hi = np.maximum(np.abs(coords[hi] - coords[idx]), F(1e-3))
        if axis == 1:
            a = np.abs((values[:, lo] - values) / gap[None, :])
            b = np.abs((values[:, hi] - values) / gap_hi[None, :])
        else:
            a = np.abs((values[lo, :] - values) / gap[:, None])
            b = np.abs((values[hi, :] - values) / gap_hi[:, None])
        return np.maximum(a, b)

    gx, gz = steps(base, left, right, 1), steps(base, left, right, 0)
    px, pz = coords[None, :], coords[:, None]
    gentle = F(1.0) - smooth(0.12, 0.3, np.hypot(gx, gz))
    mask = gentle * smooth(120.0, 400.0, np.hypot(px, pz)) * edge(px, pz)
    for cx, cz, radius in DUNE_PANS:
        mask = mask * smooth(radius + F(100.0), radius + F(450.0), np.hypot(px - cx, pz - cz))
    field = F(0.6) + F(0.6) * fbm(px / F(3_500.0), pz / F(3_500.0), seed ^ 0x51, 2)
    dunes = (small_dunes(px, pz, 260.0, F(22.0) * np.clip(field, F(0.2), F(1.0)), seed ^ 0xD00E)
             + small_dunes(px, pz, 47.0, F(2.2), seed ^ 0x7E11) * smooth(4_000.0, 1_500.0, np.hypot(px, pz)))
    return np.where(mask > F(0.0), base + mask * dunes, base).astype(F)


def erode(heights: np.ndarray, coords: np.ndarray, iterations: int) -> np.ndarray:
    """Talus relaxation between grid neighbours, conserving volume over the
    cells' unequal areas (:711-749). Each cell's changes are summed in the
    reference's loop order: from above, from the left, to the right, down."""
    h = np.asarray(heights, F).copy()
    n = coords.size
    idx = np.arange(n)
    widths = (coords[np.minimum(idx + 1, n - 1)] - coords[np.maximum(idx - 1, 0)]) * F(0.5)
    area = widths[None, :] * widths[:, None]
    inv = F(1.0) / area
    dx = (coords[1:] - coords[:-1])[None, :]
    dz = (coords[1:] - coords[:-1])[:, None]

    def transfer(delta, distance, ai, aj):
        excess = np.maximum(np.abs(delta) - F(0.65) * distance, F(0.0))
        sign = np.where(delta < F(0.0), F(-1.0), F(1.0))
        return F(0.18) * excess / (F(1.0) / ai + F(1.0) / aj) * sign

    for _ in range(iterations):
        right = transfer(h[:, :-1] - h[:, 1:], dx, area[:, :-1], area[:, 1:])    # cell i to its right
        down = transfer(h[:-1, :] - h[1:, :], dz, area[:-1, :], area[1:, :])     # cell i to the one below
        change = np.zeros_like(h)
        change[1:, :] = change[1:, :] + down / area[1:, :]
        change[:, 1:] = change[:, 1:] + right / area[:, 1:]
        change[:, :-1] = change[:, :-1] - right / area[:, :-1]
        change[:-1, :] = change[:-1, :] - down / area[:-1, :]
        h = h + change
    return h


def generate(theme: int, seed: int = SEED, surveys: Surveys | None = None) -> tuple[np.ndarray, int]:
    """``Terrain::generate``'s heights (src/terrain.rs:751-790) for a ground
    theme: ``(heights (W, W) [z, x], cells)``; ``terrain_data`` builds the
    buffer. The alpine and basalt grounds need the elevation ``surveys``."""
    cells = THEME_CELLS[theme]
    coords = coordinates(cells)
    x, z = coords[None, :], coords[:, None]
    if theme in (3, 4) and surveys is None:
        raise ValueError(f"ground theme {theme} needs the reference's elevation surveys (Surveys)")
    builders = {0: lambda: elevation(x, z, seed), 1: lambda: desert_elevation(x, z, seed),
                2: lambda: crater_elevation(x, z, seed), 3: lambda: alpine_elevation(x, z, surveys, seed),
                4: lambda: basalt_elevation(x, z, surveys, seed), 5: lambda: dune_elevation(x, z, seed)}
    heights = np.broadcast_to(builders[theme](), (cells + 1, cells + 1)).astype(F)
    if theme == 5:
        heights = superimpose_dunes(heights, coords, seed)
    if theme in (0, 1):
        heights = erode(heights, coords, 18)
    return heights, cells


def lookout_height(heights: np.ndarray, cells: int, radius_m: float) -> float:
    """``terrain::lookout_height`` (src/terrain.rs:64-71): the height under
    the lookout, metres, on a host of ``radius_m``."""
    return float(F(height(heights, cells, 0.0, 0.0)) * F(scale(radius_m)))


def stand_in(theme: int) -> tuple[np.ndarray, int]:
    """A seeded ridged-fBm heightfield at a ground's grid size, for a ground
    whose surveys are not at hand."""
    cells = THEME_CELLS[theme]
    return ridged_fbm(cells, seed=0x743337A9 ^ (theme * 0x9E3779B9 & 0xFFFFFFFF), amplitude=60.0), cells


_FOR_THEME: dict = {}


def for_theme(theme: int) -> tuple[np.ndarray, int]:
    """``terrain::for_theme`` (src/terrain.rs:58-62): each ground's terrain,
    generated once per process. The alpine and basalt grounds read the
    reference's elevation surveys from the folder ``STARGAZE_TERRAIN_SURVEYS``
    names; without it they stand on ``stand_in``."""
    import os

    if theme not in _FOR_THEME:
        folder = os.environ.get("STARGAZE_TERRAIN_SURVEYS")
        if theme in (3, 4) and not folder:
            _FOR_THEME[theme] = stand_in(theme)
        else:
            _FOR_THEME[theme] = generate(theme, surveys=Surveys(folder) if theme in (3, 4) else None)
    return _FOR_THEME[theme]
