"""Terrain traversal against exhaustive FP64 triangles.

Restates the reference's terrain oracle test (src/ground_tests.rs:740-900):
the quadtree traversal must find the same first hit as an exhaustive FP64
Moller-Trumbore scan of every grid triangle, |dt| < 0.005 + 1e-5 t, with unit
normals pointing up. Shared by the oracle (CPU) and HIP (GPU) tests.
"""
from __future__ import annotations

import numpy as np

from stargaze_amd import scenes
from stargaze_amd import terrain as tr
from stargaze_amd.layouts import PROBE_RAY

RAYS = [  # src/ground_tests.rs:763-774
    ([0.0, 300.0, 0.0], [0.0, -1.0, 0.0]),
    ([0.0, 2.0, 0.0], [0.0, 0.03, 1.0]),
    ([0.0, 2.0, 0.0], [1.0, 0.08, 0.25]),
    ([0.0, 2.0, 0.0], [-1.0, 0.1, -0.4]),
    ([0.0, 2.0, 0.0], [0.0, 1.0, 0.0]),
    ([1000.0, 2000.0, -3000.0], [0.0, -1.0, 0.0]),
    ([-4000.0, 1500.0, 2000.0], [1.0, -0.2, 0.1]),
    ([16000.0, 100.0, 0.0], [0.0, -1.0, 0.0]),
    ([16100.0, 100.0, 0.0], [0.0, -1.0, 0.0]),
    ([0.0, 0.1, 0.0], [0.5, -0.01, -1.0]),
]


def make_terrain(cells: int, seed: int = 0x743337A9):
    heights = tr.ridged_fbm(cells, seed=seed, amplitude=800.0)
    return heights, tr.terrain_data(heights, cells)


def probe_frame(cells: int, radius: float):
    f = scenes.frame()
    f.globals["ground_up"][3] = radius
    f.globals["ground_counts"] = [0, 1, 0, 0]
    return f


def probe_rays(radius: float, extra: int = 0, seed: int = 5):
    s = tr.scale(radius)
    rays = []
    for o, d in RAYS:
        rays.append((np.array(o, np.float32) * np.float32(s), np.array(d, np.float64)))
    rng = np.random.default_rng(seed)
    for _ in range(extra):
        o = np.array([rng.uniform(-3000, 3000), rng.uniform(5, 1500), rng.uniform(-3000, 3000)], np.float32) * np.float32(s)
        d = np.array([rng.normal(), -abs(rng.normal()) - 0.05, rng.normal()])
        rays.append((o, d))
    out = np.zeros(len(rays), PROBE_RAY)
    for i, (o, d) in enumerate(rays):
        out[i]["offset"] = o
        out[i]["anchor"] = 0xFFFFFFFF
        out[i]["direction"] = (d / np.linalg.norm(d)).astype(np.float32)
        out[i]["target"] = 0
    return out


def exhaustive(heights, cells, radius, rays):
    """FP64 first hit over every triangle (src/ground_tests.rs:740-760,861-900)."""
    s = float(tr.scale(radius))
    coord = tr.coordinates(cells).astype(np.float64) * s
    X, Z = np.meshgrid(coord, coord)
    r = float(radius)
    curv = -(X * X + Z * Z) / (r + np.sqrt(r * r - X * X - Z * Z))
    V = np.stack([X, heights.astype(np.float64) * s + curv, Z], -1)
    a = V[:-1, :-1].reshape(-1, 3)
    b = V[:-1, 1:].reshape(-1, 3)
    c = V[1:, :-1].reshape(-1, 3)
    e = V[1:, 1:].reshape(-1, 3)
    out = []
    for ray in rays:
        o = ray["offset"].astype(np.float64)
        d = ray["direction"].astype(np.float64)
        best = np.inf
        for (p0, p1, p2) in ((a, c, b), (e, b, c)):
            ab, ac = p1 - p0, p2 - p0
            p = np.cross(d, ac)
            det = np.einsum("ij,ij->i", ab, p)
            ok = np.abs(det) >= 1e-15
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            rel = o - p0
            u = np.einsum("ij,ij->i", rel, p) * inv
            q = np.cross(rel, ab)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", ac, q) * inv
            hit = ok & (u >= -1e-9) & (v >= -1e-9) & (u + v <= 1 + 1e-9) & (t > 0.001)
            if hit.any():
                best = min(best, float(t[hit].min()))
        out.append(best)
    return np.array(out)


def exhaustive_culled(heights, cells, radius, rays, reach=60_000.0):
    """FP64 first hit over every triangle a ray can meet, for grids too large
    to scan whole: a cell is kept when its xz rectangle lies within its
    half-diagonal of the ray's xz segment (0 to ``reach`` m), a superset of
    the cells the ray crosses, so the minimum is the exhaustive one."""
    s = float(tr.scale(radius))
    coord = tr.coordinates(cells).astype(np.float64) * s
    r = float(radius)
    h = heights.astype(np.float64) * s
    x0, x1 = coord[:-1], coord[1:]
    cx, hx = 0.5 * (x0 + x1), 0.5 * (x1 - x0)
    out = []
    for ray in rays:
        o = ray["offset"].astype(np.float64)
        d = ray["direction"].astype(np.float64)
        dxz = np.array([d[0], d[2]])
        length = np.hypot(*dxz)
        best = np.inf
        if length < 1e-12:
            # Vertical ray: only the cell below the origin.
            ix = np.clip(np.searchsorted(coord, o[0]) - 1, 0, cells - 1)
            iz = np.clip(np.searchsorted(coord, o[2]) - 1, 0, cells - 1)
            cand = [(iz, ix)]
        else:
            u = dxz / length
            span = reach * length
            # Distance from each cell centre to the segment, separably in x and z.
            px = cx[None, :] - o[0]
            pz = cx[:, None] - o[2]
            along = np.clip(px * u[0] + pz * u[1], 0.0, span)
            ex = px - along * u[0]
            ez = pz - along * u[1]
            rad = np.hypot(hx[None, :], hx[:, None])
            iz, ix = np.nonzero(ex * ex + ez * ez <= rad * rad * 1.0001)
            cand = list(zip(iz, ix))
        iz = np.array([c[0] for c in cand])
        ix = np.array([c[1] for c in cand])

        def vert(zz, xx):
            X, Z = coord[xx], coord[zz]
            curv = -(X * X + Z * Z) / (r + np.sqrt(r * r - X * X - Z * Z))
            return np.stack([X, h[zz, xx] + curv, Z], -1)

        a, b, c, e = vert(iz, ix), vert(iz, ix + 1), vert(iz + 1, ix), vert(iz + 1, ix + 1)
        for (p0, p1, p2) in ((a, c, b), (e, b, c)):
            ab, ac = p1 - p0, p2 - p0
            p = np.cross(d, ac)
            det = np.einsum("ij,ij->i", ab, p)
            ok = np.abs(det) >= 1e-15
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            rel = o - p0
            uu = np.einsum("ij,ij->i", rel, p) * inv
            q = np.cross(rel, ab)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", ac, q) * inv
            hit = ok & (uu >= -1e-9) & (v >= -1e-9) & (uu + v <= 1 + 1e-9) & (t > 0.001)
            if hit.any():
                best = min(best, float(t[hit].min()))
        out.append(best)
    return np.array(out)


def check(result, expected):
    for i, (res, exp) in enumerate(zip(result, expected)):
        actual = float(res["distance"])
        if np.isinf(exp):
            assert actual > 1e29, (i, actual)
        else:
            assert abs(actual - exp) < 0.005 + exp * 1e-5, (i, actual, exp)
            n = np.asarray(res["normal"], np.float64)
            assert np.all(np.isfinite(n)) and abs(np.linalg.norm(n) - 1.0) < 1e-4 and n[1] > 0, (i, n)
