"""Ground themes 1 (desert), 2 (lunar crater floor), 3 (alpine ice field) and
5 (dunes) on MI355X against the FP32 oracle (src/shaders/ground.wgsl:281-666),
on the Halo scene within the reference's capacities."""
import numpy as np
import pytest

import oracle
import parity
from stargaze_amd import scenes
from stargaze_amd._native import NativeError
from stargaze_amd.pathtracer import Options, PathTracer

pytestmark = pytest.mark.gpu


def themed(theme, size):
    frame, data, cells = scenes.halo_parity(*size)
    frame.globals["ground_counts"][2] = theme
    return frame, data, cells


@pytest.mark.parametrize("theme", [1, 2, 3, 5], ids=["desert", "crater", "alpine", "dunes"])
def test_ground_theme_matches_oracle_per_tile(theme):
    print(f"ground theme {theme}", flush=True)
    size = (96, 54)
    frame, data, cells = themed(theme, size)
    opts = dict(samples_per_frame=8, max_bounces=4, milky_way=0.0)
    gpu = PathTracer(size, options=Options(**opts))
    gpu.upload_terrain(data, cells)
    gpu.enable_moments(True)
    orc = oracle.OracleTracer(size, Options(**opts), "f32")
    orc.upload_terrain(data, cells)
    orc.enable_moments(True)
    for _ in range(2):
        gpu.encode(frame)
        orc.encode(frame)
    ga, gm, oa, om = gpu.read_accumulation(), gpu.read_moments(), orc.read_accumulation(), orc.read_moments()
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    agree = parity.pixel_agreement(ga, oa)
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    print(f"theme {theme}: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.5, agree
    # The theme changes the image: the forest floor renders differently.
    forest, _, _ = themed(0, size)
    plain = PathTracer(size, options=Options(**opts))
    plain.upload_terrain(data, cells)
    plain.encode(forest)
    plain.encode(forest)
    assert not np.allclose(plain.read_accumulation(), ga)


def test_unbuilt_ground_themes_are_rejected():
    size = (16, 12)
    frame, data, cells = themed(4, size)
    t = PathTracer(size, options=Options(samples_per_frame=2, max_bounces=2, milky_way=0.0))
    t.upload_terrain(data, cells)
    with pytest.raises(NativeError, match="not built"):
        t.encode(frame)
