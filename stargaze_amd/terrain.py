"""Heightfield data for the terrain quadtree, in the reference's buffer layout.

``Terrain::data`` (src/terrain.rs:47-52, built at :781-837) is
``[heights W*W | x-derivatives W*W | z-derivatives W*W | min,max per node]``
with the derivative blocks present only for grids of at most 1024 cells, and
node bounds in heap order (node k's children are 4k+1..4k+4, child c offset
by (c & 1, c >> 1) half-tiles). The HIP library turns this into cached,
curved AABBs and vertices on the GPU (cache_terrain_geometry).

The heights themselves come from a seeded ridged fBm generator written for
the benchmark workloads; the reference's own generator (warped multifractal
noise with erosion, src/terrain.rs:83-780) is not restated.
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
