"""The product's stackless traversals, built for the host, against the
reference's flat primitive loop and the oracle (CPU)."""
import subprocess
import sys
from pathlib import Path

import numpy as np

from stargaze_amd.layouts import MISS

HERE = Path(__file__).resolve().parent


def test_bvh_and_flat_loop_give_the_same_landscape_answers(tmp_path):
    out = tmp_path / "answers.npz"
    subprocess.run([sys.executable, str(HERE / "native" / "traversal_run.py"), str(out)], check=True, timeout=300)
    r = np.load(out)
    bvh, flat = r["bvh"], r["flat"]
    occl = r["occlusion"] != 0
    closest = ~occl
    # Closest hits: same primitive and distance with and without the BVH.
    assert np.array_equal(bvh["index"][closest], flat["index"][closest])
    assert np.array_equal(bvh["distance"][closest], flat["distance"][closest])
    # Occlusion: same verdict (the first occluder found may differ).
    assert np.array_equal(bvh["index"][occl] != MISS, flat["index"][occl] != MISS)
    # And the flat loop agrees with the FP32 oracle's. Near the levelled
    # lookout the terrain triangles lie on the host sphere, so which of the
    # two is "first" is a tie that FP64 (product) and f32 (oracle) break
    # differently; those rays must still agree on the distance.
    fi, ri = flat["index"][closest], r["ref_index"][closest].astype(np.uint32)
    fd, rd = flat["distance"][closest].astype(np.float64), r["ref_distance"][closest]
    assert np.array_equal(fi == MISS, ri == MISS)
    tie = (fi != ri) & np.isin(fi, [0x80000000, 0x80100000]) & np.isin(ri, [0x80000000, 0x80100000])
    assert np.all((fi == ri) | tie)
    hit = fi != MISS
    rel = np.abs(fd[hit] - rd[hit]) / rd[hit]
    assert np.median(rel) < 1e-5 and np.all(rel[tie[hit]] < 1e-5)


import pytest  # noqa: E402


@pytest.mark.parametrize("cells,radius", [(64, 6_371_000.0), (64, 5_000.0), (128, 1_200_000.0)])
def test_terrain_traversal_matches_exhaustive_triangles(tmp_path, cells, radius):
    import terrain_cases as tc

    out = tmp_path / "terrain.npz"
    subprocess.run([sys.executable, str(HERE / "native" / "terrain_run.py"), str(out), str(cells), str(radius)],
                   check=True, timeout=300)
    r = np.load(out)
    tc.check(r["result"], r["expected"])
