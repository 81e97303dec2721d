"""The headless render request (stargaze_amd.headless) against the
reference's checks and presets (src/headless.rs:34-121,
src/graphics.rs:46-91)."""
from __future__ import annotations

import json

import pytest

from stargaze_amd import headless as hl
from stargaze_amd import system as sy

GOOD = {"version": 1, "scene": "vesper.yaml", "mode": "balanced", "width": 1280, "height": 720, "samples": 64,
        "exposure": 1.5}


def request(**change):
    return hl.Request.parse(json.dumps({**GOOD, **change}))


def test_presets_and_trace_sizes():
    assert hl.PRESETS["balanced"].render_size((1280, 720)) == (960, 540)
    assert hl.PRESETS["eco"].render_size((1280, 720)) == (640, 360)
    assert hl.PRESETS["high"].render_size((33, 19)) == (33, 19)
    assert hl.PRESETS["eco"].render_size((33, 19)) == (17, 10)
    assert [p.max_bounces for p in hl.PRESETS.values()] == [2, 4, 8]
    assert [p.vegetation_percent for p in hl.PRESETS.values()] == [35, 70, 100]


def test_request_checks():
    request().validate()
    for change, message in (({"version": 2}, "version 2"), ({"width": 15}, "dimensions"),
                            ({"height": 2161}, "dimensions"), ({"samples": 0}, "samples"),
                            ({"samples": 4097}, "samples"), ({"exposure": 0.0}, "exposure"),
                            ({"exposure": -1.0}, "exposure")):
        with pytest.raises(hl.RequestError, match=message):
            request(**change).validate()
    for bad in ({**GOOD, "mode": "ultra"}, {**GOOD, "colour": 1}, {k: v for k, v in GOOD.items() if k != "scene"}):
        with pytest.raises(hl.RequestError):
            hl.Request.parse(json.dumps(bad))


def test_scene_is_relative_to_the_request_and_needs_a_time(tmp_path):
    (tmp_path / "vesper.yaml").write_text((sy.CONFIGS / "vesper.yaml").read_text())
    (tmp_path / "halo.yaml").write_text((sy.CONFIGS / "halo.yaml").read_text())
    scene = hl.load_scene(request(), tmp_path / "request.json")
    assert scene.initial_time_days == 0.0
    with pytest.raises(hl.RequestError, match="explicit time_days"):
        hl.load_scene(request(scene="halo.yaml"), tmp_path / "request.json")


def test_command_line_arguments(tmp_path):
    for argv, message in ((["--nope"], "unknown render argument"), (["--request"], "missing"),
                          (["--request", "a", "--request", "b"], "duplicate"), (["--output", "x.png"], "--request"),
                          (["--request", "a"], "--output")):
        with pytest.raises(hl.RequestError, match=message):
            hl.run(argv)
    assert hl.run(["--help"]) == 0
