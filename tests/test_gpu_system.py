"""Frames assembled from the reference's scene files (stargaze_amd.system)
rendered by the HIP path against the FP32 oracle per 16x16 tile: two suns
beside a ringed giant, a ringed star among styled moons over desert rocks,
and the Moon through a 1.1 degree telescope over the forest clearing.

The level-one file asks for the basalt ground, whose materials need the
reference's volcano survey data and are not built (the library and the
oracle refuse theme 4); its frame is rendered over the desert ground."""
import numpy as np
import pytest

import oracle
import parity
from stargaze_amd import system as sy
from stargaze_amd import terrain as tr
from stargaze_amd.pathtracer import Options, PathTracer
from test_system import text

pytestmark = pytest.mark.gpu

# (scene file, time in days or None for the file's own, terrain cells, ground override)
FILES = [("1-halo-volcano.yaml", None, 128, "desert"), ("corona-valley.yaml", None, 128, None),
         ("solar-system.yaml", 17.25, 128, None)]


@pytest.mark.parametrize("name,time,cells,ground", FILES, ids=[f[0] for f in FILES])
def test_scene_file_frame_matches_oracle_per_tile(name, time, cells, ground):
    scene = sy.parse(text(name))
    if ground is not None:
        scene.ground = sy.THEMES[ground]
    t = scene.initial_time_days if time is None else time
    size = (96, 54)
    terrain = (tr.ridged_fbm(cells, seed=0x743337A9, amplitude=60.0), cells) if cells else None
    frame = sy.build_frame(scene, t, *size, terrain=terrain)
    print(f"{name} at day {t}: {len(frame.bodies)} bodies, {len(frame.ground)} props", flush=True)
    opts = dict(samples_per_frame=8, max_bounces=4, milky_way=0.0)
    gpu = PathTracer(size, options=Options(**opts))
    orc = oracle.OracleTracer(size, Options(**opts), "f32")
    if terrain is not None:
        data = tr.terrain_data(*terrain)
        gpu.upload_terrain(data, cells)
        orc.upload_terrain(data, cells)
    gpu.enable_moments(True)
    orc.enable_moments(True)
    for _ in range(2):
        gpu.encode(frame)
        orc.encode(frame)
    ga, gm, oa, om = gpu.read_accumulation(), gpu.read_moments(), orc.read_accumulation(), orc.read_moments()
    gpu.close()
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    assert (oa[..., :3].sum(-1) > 0).mean() > 0.1, "the view is mostly black"
    agree = parity.pixel_agreement(ga, oa)
    over, worst = parity.assert_tiles(parity.tile_z(ga, gm, oa, om))
    print(f"{name}: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.99, agree
