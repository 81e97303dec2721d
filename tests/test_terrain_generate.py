"""The reference's terrain generator (stargaze_amd.terrain.generate) against
the reference's own terrain tests (src/terrain.rs:870-995). The alpine and
basalt grounds read the reference's elevation surveys; their checks run
where a folder of them is named by STARGAZE_TERRAIN_SURVEYS (or the
reference checkout is present) and are skipped otherwise."""
from __future__ import annotations

import os
import pathlib
import re

import numpy as np
import pytest

from stargaze_amd import terrain as tr

SURVEYS = pathlib.Path(os.environ.get("STARGAZE_TERRAIN_SURVEYS", "/root/reference/assets/terrain"))
needs_surveys = pytest.mark.skipif(not (SURVEYS / "stromboli-2023.u16").exists(),
                                   reason="the reference's elevation surveys are not available here")


def _generated(theme, surveys=None):
    heights, cells = tr.generate(theme, surveys=surveys)
    return heights, cells, tr.terrain_data(heights, cells)


def _check_bounds(data, heights, cells):
    """Every node's min/max encloses its vertices (heap order, children
    4k+1..4k+4 at (c & 1, c >> 1) half-tiles)."""
    offset = tr.bounds_offset(cells)

    def check(node, x, z, size):
        lo, hi = data[offset + 2 * node], data[offset + 2 * node + 1]
        block = heights[z:z + size + 1, x:x + size + 1]
        assert block.min() >= lo and block.max() <= hi
        if size > tr.LEAF and size >= cells // 8:
            for corner in range(4):
                check(node * 4 + corner + 1, x + (corner & 1) * size // 2, z + (corner >> 1) * size // 2, size // 2)

    check(0, 0, 0, cells)


@pytest.mark.parametrize("theme", [0, 1], ids=["forest", "desert"])
def test_generated_surface_is_finite_continuous_and_bounds_enclose_every_vertex(theme):
    """src/terrain.rs:874-912."""
    heights, cells, data = _generated(theme)
    assert cells == 512 and data.size == tr.bounds_offset(cells) + 2 * tr.node_count(cells)
    assert np.all(np.isfinite(data))
    assert tr.height(heights, cells, 0.0, 0.0) == 0.0
    assert heights.min() >= -1e-5 and heights.max() > 500.0
    _check_bounds(data, heights, cells)
    for border in (heights[0], heights[cells], heights[:, 0], heights[:, cells]):
        assert np.abs(border).max() < 1e-6
    for x, z in ((100.0, 230.0), (2_000.0, -3_000.0), (-4_000.0, 1_000.0)):
        assert abs(tr.height(heights, cells, x + 0.001, z) - tr.height(heights, cells, x - 0.001, z)) < 0.1


def _explore(theme, surveys=None):
    heights, cells, data = _generated(theme, surveys)
    assert data.size == tr.bounds_offset(cells) + 2 * tr.node_count(cells)
    assert np.all(np.isfinite(data)) and heights.min() >= -1e-5
    root = tr.bounds_offset(cells)
    assert heights.min() >= data[root] and heights.max() <= data[root + 1]
    assert np.abs(heights[0]).max() < 1e-6 and np.abs(heights[:, cells]).max() < 1e-6
    for x, z in ((100.0, 230.0), (2_000.0, -3_000.0)):
        assert abs(tr.height(heights, cells, x + 0.001, z) - tr.height(heights, cells, x - 0.001, z)) < 0.1
    return heights, cells, data


def test_explore_grounds_are_finite_bounded_and_meet_the_sphere_at_the_edge():
    """src/terrain.rs:913-949 for the crater and dune grounds: the crater
    plain and the dune crest lift the eye onto the surface."""
    heights, cells, _ = _explore(2)
    assert cells == tr.CRATER_CELLS and tr.lookout_height(heights, cells, 6.371e6) > 20.0
    heights, cells, _ = _explore(5)
    assert cells == 1024 and tr.lookout_height(heights, cells, 6.371e6) > 300.0


@needs_surveys
def test_surveyed_grounds_are_finite_bounded_and_lift_the_eye():
    """src/terrain.rs:913-949 for the alpine and basalt grounds: the nunatak
    and the volcano summit lift the eye, and the ice field's peaks pass
    1500 m."""
    surveys = tr.Surveys(SURVEYS)
    heights, cells, data = _explore(3, surveys)
    assert cells == 1024 and tr.lookout_height(heights, cells, 6.371e6) > 150.0
    assert data[tr.bounds_offset(cells) + 1] > 1_500.0
    heights, cells, _ = _explore(4, surveys)
    assert cells == tr.MAX_CELLS and tr.lookout_height(heights, cells, 6.371e6) > 200.0


def test_erosion_moves_material_downhill_and_conserves_volume():
    """src/terrain.rs:973-994."""
    coords = np.array([0.0, 1.0, 2.0, 3.0, 4.0], np.float32)
    heights = np.zeros((5, 5), np.float32)
    heights[2, 2] = 10.0
    out = tr.erode(heights, coords, 12)
    assert out[2, 2] < 10.0 and out[2, 1] > 0.0 and out.min() >= 0.0
    w = np.array([0.5, 1.0, 1.0, 1.0, 0.5])
    assert abs(float((out * w[None, :] * w[:, None]).sum()) - 10.0) < 1e-4
    assert tr.elevation(np.float32(1_230.0), np.float32(4_560.0), tr.SEED) != \
        tr.elevation(np.float32(1_230.0), np.float32(4_560.0), tr.SEED + 1)


def test_erosion_sums_each_cell_in_loop_order():
    """The vectorised relaxation equals a literal transcription of the
    reference's loop (cell by cell, right then down, one change buffer) bit
    for bit on a random 9 x 9 patch with uneven spacing."""
    rng = np.random.default_rng(3)
    coords = np.cumsum(rng.uniform(0.5, 3.0, 9)).astype(np.float32)
    heights = rng.uniform(0.0, 20.0, (9, 9)).astype(np.float32)
    f = np.float32
    n = 9
    widths = [(coords[min(i + 1, n - 1)] - coords[max(i - 1, 0)]) * f(0.5) for i in range(n)]
    area = [widths[i % n] * widths[i // n] for i in range(n * n)]
    h = heights.reshape(-1).copy()
    for _ in range(3):
        change = np.zeros(n * n, np.float32)
        for z in range(n):
            for x in range(n):
                i = z * n + x
                for j, distance in ((z * n + min(x + 1, n - 1), coords[min(x + 1, n - 1)] - coords[x]),
                                    (min(z + 1, n - 1) * n + x, coords[min(z + 1, n - 1)] - coords[z])):
                    if i == j:
                        continue
                    delta = h[i] - h[j]
                    excess = max(abs(delta) - f(0.65) * distance, f(0.0))
                    volume = f(0.18) * excess / (f(1.0) / area[i] + f(1.0) / area[j]) * (f(-1.0) if delta < 0 else f(1.0))
                    change[i] = change[i] - volume / area[i]
                    change[j] = change[j] + volume / area[j]
        h = h + change
    np.testing.assert_array_equal(tr.erode(heights, coords, 3).reshape(-1), h)


def test_kernel_materials_mirror_the_terrain_constants():
    """src/terrain.rs:951-971: the ground materials carry the generator's
    dune wind, pan level and pans, ice flow and nunatak; here the HIP
    kernels' and the oracle's copies against this module's."""
    root = pathlib.Path(tr.__file__).resolve().parent
    kernel = (root / "csrc" / "sg_integrator.h").read_text()
    oracle = (root.parent / "oracle" / "sg_oracle.c").read_text()

    def constant(text, name):
        m = re.search(rf"{name}\s*=?\s*([0-9.]+)f?", text)
        return float(m.group(1))

    for text in (kernel, oracle):
        assert constant(text, "ICE_FLOW") == pytest.approx(float(tr.ICE_FLOW), abs=1e-6)
        assert constant(text, "DUNE_WIND") == pytest.approx(float(tr.DUNE_WIND), abs=1e-6)
        assert constant(text, "DUNE_PAN_LEVEL") == float(tr.DUNE_PAN_LEVEL)
        pans = [float(v) for v in re.findall(r"pans\[2\]\[3\] = \{\{([^}]*)\}, \{([^}]*)\}\}", text)[0][0]
                .replace("f", "").split(",")] + \
               [float(v) for v in re.findall(r"pans\[2\]\[3\] = \{\{([^}]*)\}, \{([^}]*)\}\}", text)[0][1]
                .replace("f", "").split(",")]
        np.testing.assert_allclose(pans, np.array(tr.DUNE_PANS, np.float64).reshape(-1), rtol=1e-6)
    assert "/ 700.0f, 1.0f)" in kernel and "return 180.0f *" in kernel
    r = np.array([0.0, 100.0, 400.0, 800.0], np.float32)
    u = np.minimum(np.sqrt(r * r + 900.0) / 700.0, 1.0)
    np.testing.assert_allclose(tr.lookout_hill(r), 180.0 * (1 - u) ** 2, rtol=1e-6)
