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


def test_system_views_follow_the_batched_config():
    a = scenes.system_views(24, seed=1, min_pixels=32 * 32, max_pixels=160 * 90)
    b = scenes.system_views(24, seed=1, min_pixels=32 * 32, max_pixels=160 * 90)
    for v, w in zip(a, b):
        assert v["frame"].bodies.tobytes() == w["frame"].bodies.tobytes()
        assert v["spp"] in (4, 8, 16, 32, 64)
        assert 2 <= v["bodies"] <= 12 and len(v["frame"].bodies) == v["bodies"]
        assert v["frame"].bodies["emissive"][0] == 1.0 and v["frame"].bodies["emissive"][1] == 0.0
        g = v["frame"].globals
        assert (g["atmo_params"][3] >= 1.0) == (v["kind"] == "air")
        # camera 2 m above the host
        b1 = v["frame"].bodies[1]
        centre = b1["center"].astype(np.float64) + b1["center_low"][:3].astype(np.float64)
        alt = np.linalg.norm(centre) - float(b1["radius"])
        assert abs(alt * scenes.METRES_PER_AU - 2.0) < 0.5  # f32 radius: 6e-8 relative of 1e6 m
    kinds = {v["kind"] for v in scenes.system_views(96)}
    assert kinds == {"air", "airless"}
