"""Terrain data layout and the oracle's quadtree traversal (CPU)."""
import numpy as np
import pytest

import oracle
import terrain_cases as tc
from stargaze_amd import terrain as tr
from stargaze_amd.pathtracer import Options


@pytest.mark.parametrize("cells", [8, 64, 2048])
def test_node_bounds_enclose_their_vertices(cells):
    """src/terrain.rs:870-992: every quadtree node bounds the vertices of its tile."""
    heights = tr.ridged_fbm(cells, amplitude=800.0) if cells <= 64 else np.random.default_rng(1).normal(
        size=(cells + 1, cells + 1)).astype(np.float32) * 50
    data = tr.terrain_data(heights, cells)
    off = tr.bounds_offset(cells)
    rng = np.random.default_rng(0)
    nodes = tr.node_count(cells)
    for k in list(range(min(nodes, 200))) + list(rng.integers(0, nodes, 300)):
        k = int(k)
        level = int(np.floor(np.log2(3 * k + 1))) // 2
        size = cells >> level
        m = k - (4 ** level - 1) // 3
        x = int(''.join(reversed(bin(m)[2:].zfill(32)))[0::2][::-1], 2) * size
        z = int(''.join(reversed(bin(m)[2:].zfill(32)))[1::2][::-1], 2) * size
        tile = heights[z: z + size + 1, x: x + size + 1]
        assert data[off + 2 * k] == tile.min() and data[off + 2 * k + 1] == tile.max()


@pytest.mark.parametrize("radius", [6_371_000.0, 5_000.0])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_oracle_terrain_matches_exhaustive_triangles(radius, precision):
    cells = 64
    heights, data = tc.make_terrain(cells)
    rays = tc.probe_rays(radius, extra=40)
    t = oracle.OracleTracer((4, 4), Options(milky_way=0.0), precision)
    t.upload_terrain(data, cells)
    res = t.probe(tc.probe_frame(cells, radius), 4, rays)
    tc.check(res, tc.exhaustive(heights, cells, radius, rays))
