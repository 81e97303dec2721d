"""The reference's render benchmark cases (stargaze_amd.render_benchmark)
against the reference's own assertions on them
(src/render_benchmark.rs:999-1182), plus the run configuration's checks, the
fingerprint and the star catalogue."""
from __future__ import annotations

import json
import math
import struct
import zlib

import numpy as np
import pytest

from stargaze_amd import render_benchmark as rb
from stargaze_amd import stars as st
from stargaze_amd.layouts import AU_KM

CFG = rb.Config(output="", width=1280, height=720, scale=75, bounces=4, vegetation=70, warmup=3, frames=8,
                repeats=3, image_samples=32)


def fixture(name, step=0, cfg=CFG):
    return rb.fixture(name, step, cfg)


def centre(frame, i) -> np.ndarray:
    return frame.bodies[i]["center"].astype(np.float64)


def up_of(frame) -> np.ndarray:
    return frame.globals["ground_up"][:3].astype(np.float32)


def sun_record(frame):
    return next(b for b in frame.bodies if b["emissive"] > 0.5)


def test_config_checks_and_trace_size():
    """src/render_benchmark.rs:198-250."""
    CFG.validate()
    assert CFG.trace_size() == (960, 540)
    assert rb.Config(**{**CFG.__dict__, "width": 1001, "scale": 33}).trace_size() == (331, 238)
    for change, message in (({"variant": "guided-ground"}, "variant"), ({"width": 15}, "dimensions"),
                            ({"height": 2161}, "dimensions"), ({"scale": 24}, "quality"), ({"bounces": 17}, "quality"),
                            ({"vegetation": 101}, "vegetation"), ({"warmup": 0}, "warmup"), ({"frames": 1}, "frames"),
                            ({"repeats": 21}, "repeats"), ({"image_samples": 1}, "image samples"),
                            ({"motion_reference_samples": 1}, "motion"), ({"cases": []}, "case"),
                            ({"cases": ["nowhere"]}, "case"), ({"cases": ["level-2", "level-2"]}, "duplicate")):
        with pytest.raises(ValueError, match=message):
            rb.Config(**{**CFG.__dict__, **change}).validate()
    text = json.dumps({**CFG.__dict__, "cases": ["level-2"]})
    assert rb.Config.from_json(text).cases == ["level-2"]
    with pytest.raises(ValueError, match="unknown config"):
        rb.Config.from_json(json.dumps({**CFG.__dict__, "colour": 1}))


def test_fingerprint_is_fnv1a_of_the_frame_bytes():
    """FNV-1a 64 of the empty input is its offset basis; of "a" the published
    0xaf63dc4c8601ec8c. The frame's fingerprint covers globals, bodies,
    ground and time."""
    def fnv(data: bytes) -> str:
        h = 0xCBF29CE484222325
        for b in data:
            h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        return f"{h:016x}"

    assert fnv(b"") == "cbf29ce484222325" and fnv(b"a") == "af63dc4c8601ec8c"
    f = fixture("dual-eclipse")
    data = f.globals.tobytes() + f.bodies.tobytes() + f.ground.tobytes() + struct.pack("<d", f.scene_time)
    assert rb.fingerprint(f) == fnv(data)


def test_fixtures_are_deterministic_and_cover_requested_workloads():
    """src/render_benchmark.rs:999-1182, every case at 1280x720 with
    vegetation 70, on a levelled ground."""
    for name in rb.CASES:
        a = fixture(name)
        assert rb.fingerprint(a) == rb.fingerprint(fixture(name))
        assert rb.fingerprint(a) != rb.fingerprint(fixture(name, 1))
        if name.startswith("atacama-"):
            assert a.globals["ground_counts"][2] == 1 and len(a.ground) > 0
            s = sun_record(a)
            altitude = s["center"] / np.linalg.norm(s["center"]) @ up_of(a)
            assert (altitude > 0.0) == (name == "atacama-day")
        assert (a.globals["atmo_params"][3] > 0.0) == (name not in ("vantus-airless", "dual-eclipse",
                                                                     "ground-craters")), name
        if name in ("earth-daylight", "earth-twilight", "median-dense"):
            s = a.bodies[1] if name == "median-dense" else sun_record(a)
            altitude = s["center"] / np.linalg.norm(s["center"]) @ up_of(a)
            if name == "earth-twilight":
                assert -0.05 <= altitude < 0.0
            else:
                assert altitude > 0.2, f"{name} must be daylight"
            assert len(a.ground) > 0
        if name == "twilight-afterglow":
            up = up_of(a).astype(np.float64)
            s = centre(a, next(i for i, b in enumerate(a.bodies) if b["emissive"] > 0.5))
            s /= np.linalg.norm(s)
            assert -6.0 <= math.degrees(math.asin(s @ up)) < -4.0
            flat = lambda v: (v - up * (v @ up)) / np.linalg.norm(v - up * (v @ up))  # noqa: E731
            offset = math.degrees(math.atan2(flat(s) @ flat(np.array([1.0, 0, 0])),
                                             flat(s) @ flat(np.array([0, 0, -1.0]))))
            assert 20.0 <= offset < 30.0, offset
            assert len(a.ground) > 0
        if name == "dual-eclipse":
            assert sum(b["emissive"] > 0.5 for b in a.bodies) == 2
            for sun, blocker in ((1, 3), (2, 4)):
                midpoint = (np.array([0.0, 0.0, -2.0], np.float32) + a.bodies[sun]["center"]) * 0.5
                assert np.linalg.norm(midpoint - a.bodies[blocker]["center"]) < 1e-6


def _discs(frame, point):
    sun, moon = centre(frame, 1) - point, centre(frame, 2) - point
    separation = math.asin(np.linalg.norm(np.cross(sun / np.linalg.norm(sun), moon / np.linalg.norm(moon))))
    return (separation, math.asin(float(frame.bodies[1]["radius"]) / np.linalg.norm(sun)),
            math.asin(float(frame.bodies[2]["radius"]) / np.linalg.norm(moon)))


def test_umbra_edge_geometry():
    """src/render_benchmark.rs:1053-1102: the observer and the right-hand air
    in totality, the left-hand air seeing the whole star, the star unobscured
    40 s earlier, its disc far below a pixel, and the sunlit sample in dense
    air."""
    km = 1.0 / AU_KM
    for step in (0, 8, 15):
        f = fixture("eclipse-umbra-edge", step)
        for point in (np.zeros(3), np.array([15.0, 5.0, -30.0]) * km):
            s, sun, moon = _discs(f, point)
            assert s + sun < moon
        s, sun, moon = _discs(f, np.array([-15.0, 5.0, -30.0]) * km)
        assert s > sun + moon
    a = fixture("eclipse-umbra-edge")
    before = fixture("eclipse-umbra-edge")
    before.bodies[2]["center"][0] += np.float32(4.0 * km)
    s, sun, moon = _discs(before, np.zeros(3))
    assert s > sun + moon
    _, sun, _ = _discs(a, np.zeros(3))
    assert math.tan(sun) * CFG.trace_size()[1] / float(a.globals["cam_forward"][3]) < 0.02
    host = a.globals["atmo_center"][:3].astype(np.float64)
    height = np.linalg.norm(np.array([-15.0, 5.0, -30.0]) * km - host) - float(a.globals["atmo_center"][3])
    assert 5.0 * km <= height < 20.0 * km


def test_totality_geometry():
    """src/render_benchmark.rs:1125-1168: the sun 60 degrees up, the observer
    and air 20 km away in totality, low air 300 km away in the penumbra."""
    km = 1.0 / AU_KM
    a = fixture("eclipse-totality")
    up = up_of(a).astype(np.float64)
    sun = centre(a, 1) / np.linalg.norm(centre(a, 1))
    assert abs(math.degrees(math.asin(sun @ up)) - 60.0) < 1e-3
    for step in (0, 8, 15):
        f = fixture("eclipse-totality", step)
        for point in (np.zeros(3), np.array([20.0, 0, 0]) * km, np.array([-20.0, 0, 0]) * km):
            s, r_sun, r_moon = _discs(f, point)
            assert s + r_sun < r_moon, f"step {step}: must be in totality"
        for azimuth in (0.0, 90.0, 180.0, 270.0):
            t = math.radians(azimuth)
            p = np.array([math.sin(t), 0.0, -math.cos(t)]) * (300.0 * km) + up * (5.0 * km)
            s, r_sun, r_moon = _discs(f, p)
            assert s + r_sun > r_moon and s < r_sun + r_moon


def test_level_and_ground_cases_use_their_grounds_and_lift_the_eye():
    """The level cases render their files' grounds; the ground cases the
    explore grounds at the Atacama site; a heightfield lifts the eye by its
    height under the lookout (terrain::lookout_height)."""
    themes = {"level-1": 4, "level-2": 5, "level-3": 3, "ground-craters": 2, "ground-alpine": 3,
              "ground-basalt": 4, "ground-dunes": 5}
    for name, theme in themes.items():
        assert fixture(name).globals["ground_counts"][2] == theme
        assert rb.ground_theme(name) == theme
    heights, cells = rb.stand_in_terrain(5)
    lifted = rb.fixture("ground-dunes", 0, CFG, (heights, cells))
    flat = fixture("ground-dunes")
    from stargaze_amd import system as sy
    scene = sy.load(sy.CONFIGS / "atacama.yaml")
    lay = sy.surface_height((heights, cells), scene.bodies[scene.host].radius * sy.METRES_PER_AU)(0.0, 0.0)
    assert abs(lifted.globals["ground_east"][3] - flat.globals["ground_east"][3] - lay) < 1e-3


def test_star_catalogue():
    """stars::generate (src/stars.rs:91-115): 3500 + 2500 stars from one
    stream, unit directions, the power-law brightness and disc radius ranges,
    the band about its great circle, and Tanner Helland's fit at its fixed
    points (white at 6600 K; red saturated below it, blue above)."""
    cat = st.generate(3500, 2500, 0x5EED1234)
    assert len(cat) == 6000
    again = st.generate(3500, 2500, 0x5EED1234)
    assert all(a.dir == b.dir and a.color == b.color for a, b in zip(cat, again))
    dirs = np.array([c.dir for c in cat])
    assert np.abs(np.linalg.norm(dirs, axis=1) - 1.0).max() < 1e-6
    bright = np.array([c.bright for c in cat])
    assert bright[:3500].min() >= np.float32(0.02) and bright[:3500].max() <= 1.62 + 1e-6
    assert bright[3500:].max() <= 1.62 * 0.28 + 1e-6
    radius = np.array([c.angular_radius for c in cat])
    assert radius.min() >= 0.7 * 0.0004 - 1e-9 and radius.max() <= 3.4 * 0.0004 + 1e-9
    n = np.array([0.35, 0.82, 0.45]) / np.linalg.norm([0.35, 0.82, 0.45])
    # Spread is at most 0.42 out of the unit circle's plane.
    assert np.abs(dirs[3500:] @ n).max() <= 0.42 / math.hypot(1.0, 0.0) + 1e-6
    assert np.abs(dirs[:3500] @ n).max() > 0.9  # the uniform stars cover the poles of the band
    np.testing.assert_allclose(st.blackbody_linear(6600.0), (1.0, 1.0, 1.0), atol=1e-6)
    for t in (1000.0, 3000.0, 6000.0):
        assert st.blackbody_linear(t)[0] == 1.0
    for t in (7000.0, 12000.0, 40000.0):
        assert st.blackbody_linear(t)[2] == 1.0
    assert st.blackbody_linear(1500.0)[2] == 0.0


def test_png_writer_round_trips():
    import pathlib
    import tempfile

    rgba = np.random.default_rng(1).integers(0, 256, (5, 7, 4), dtype=np.uint8)

    with tempfile.TemporaryDirectory() as d:
        p = pathlib.Path(d) / "x.png"
        rb.write_png(p, rgba)
        data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h = struct.unpack(">II", data[16:24])
    assert (w, h) == (7, 5)
    start = data.index(b"IDAT") + 4
    length = struct.unpack(">I", data[start - 8:start - 4])[0]
    raw = zlib.decompress(data[start:start + length])
    rows = np.frombuffer(raw, np.uint8).reshape(5, 1 + 7 * 4)
    assert np.all(rows[:, 0] == 0)
    np.testing.assert_array_equal(rows[:, 1:].reshape(5, 7, 4), rgba)
