"""The reference's render benchmark suite run on the HIP path: every case
times, captures and lands in ``results.json`` in the reference's schema, and
the benchmark's own views render as the FP32 oracle renders them (per 16x16
tile)."""
import json

import numpy as np
import pytest

import oracle
import parity
from stargaze_amd import render_benchmark as rb
from stargaze_amd import terrain as tr
from stargaze_amd.pathtracer import Options, PathTracer

pytestmark = pytest.mark.gpu


def test_suite_runs_every_case_and_writes_results(tmp_path):
    cfg = rb.Config(output=str(tmp_path), width=48, height=27, scale=100, bounces=2, vegetation=70, warmup=1,
                    frames=2, repeats=1, image_samples=2)
    result = rb.run(cfg, log=lambda line: print(line, flush=True))
    saved = json.loads((tmp_path / "results.json").read_text())
    assert saved["suite_version"] == 1 and saved["variant"] == "reference" and saved["trace_size"] == [48, 27]
    assert [c["name"] for c in saved["cases"]] == list(rb.CASES)
    for case in saved["cases"]:
        if case["name"] in rb.NOT_BUILT:
            assert case["skipped"] == rb.NOT_BUILT[case["name"]]
            continue
        assert len(case["frame_fingerprints"]) == 2
        assert [m["mode"] for m in case["modes"]] == ["paused", "advancing"]
        paused, moving = (m["repeats"][0] for m in case["modes"])
        assert [t["spp"] for t in paused] == [1, 2] and [t["spp"] for t in moving] == [1, 1]
        assert all(t["gpu_ms"] > 0.0 and t["wall_ms"] >= t["gpu_ms"] * 0.5 for t in paused + moving)
        names = [i["name"] for i in case["images"]]
        assert names == [f"{case['name']}-spp1", f"{case['name']}-spp2", f"{case['name']}-moving-1"]
        for image in names:
            assert (tmp_path / f"{image}.png").stat().st_size > 0
            assert (tmp_path / f"{image}.rgba").stat().st_size == 48 * 27 * 4
            linear = np.frombuffer((tmp_path / f"{image}.f32").read_bytes(), np.float32)
            assert linear.size == 48 * 27 * 4 and np.all(np.isfinite(linear)) and np.all(linear >= 0)
    assert result["cases"][0]["name"] == "halo-rings"


CASES = ["dual-eclipse", "moonlit-air", "eclipse-umbra-edge", "eclipse-totality", "vantus-eclipse",
         "atacama-day", "ground-alpine", "level-3"]


@pytest.mark.parametrize("name", CASES)
def test_benchmark_view_matches_oracle_per_tile(name):
    size = (64, 36)
    cfg = rb.Config(output="", width=size[0], height=size[1], scale=100, bounces=4, vegetation=70, warmup=1,
                    frames=2, repeats=1, image_samples=2)
    theme = rb.ground_theme(name)
    terrain = rb.reference_terrain(theme) if theme is not None else None
    frame = rb.fixture(name, 0, cfg, terrain)
    print(f"{name}: {len(frame.bodies)} bodies, {len(frame.ground)} props", flush=True)
    opts = dict(samples_per_frame=8, max_bounces=4, milky_way=0.0)
    gpu = PathTracer(size, options=Options(**opts))
    orc = oracle.OracleTracer(size, Options(**opts), "f32")
    if terrain is not None:
        data = tr.terrain_data(*terrain)
        gpu.upload_terrain(data, terrain[1])
        orc.upload_terrain(data, terrain[1])
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
