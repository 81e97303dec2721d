"""The batched-view workload builder (host only)."""
import numpy as np

from stargaze_amd import scenes


def test_batch_views_shapes_and_determinism():
    a = scenes.batch_views(12, seed=4, min_pixels=16 * 16, max_pixels=200 * 120)
    b = scenes.batch_views(12, seed=4, min_pixels=16 * 16, max_pixels=200 * 120)
    assert len(a) == 12
    kinds = {v["kind"] for v in a}
    assert kinds == {"triple", "giant", "orbit_air", "halo"}
    for v, w in zip(a, b):
        assert (v["width"], v["height"], v["spp"]) == (w["width"], w["height"], w["spp"])
        assert v["frame"].globals.tobytes() == w["frame"].globals.tobytes()
        px = v["width"] * v["height"]
        assert 8 * 8 <= px <= 1.2 * 200 * 120
        assert v["spp"] in (1, 2, 4, 8, 16)
        g = v["frame"].globals
        assert g["viewport"][0] == v["width"] and g["viewport"][1] == v["height"]
        assert (v["terrain"] is not None) == (v["kind"] == "halo")
        if v["kind"] in ("halo", "orbit_air"):
            assert g["atmo_params"][3] >= 1.0
    sizes = np.array([v["width"] * v["height"] for v in scenes.batch_views(96, min_pixels=128 * 128)])
    assert sizes.min() >= 0.9 * 128 * 128 and sizes.max() <= 1.1 * 1920 * 1080
