"""Procedural surface styles 0-6 on MI355X against the FP32 oracle: the
material probe (src/surface_tests.rs:87-169) and rendered scenes."""
import numpy as np
import pytest

import oracle
import parity
from stargaze_amd import scenes
from stargaze_amd._native import NativeError
from stargaze_amd.pathtracer import PROBE_ALBEDO, Options, PathTracer

pytestmark = pytest.mark.gpu

STYLES = tuple(range(7))  # plain, gas giant, cratered, rocky, icy, volcanic, terran


def style_frame(style):
    f = scenes.material_frame(scenes.STYLE_PLAIN if style == scenes.STYLE_GAS_GIANT else style)
    if style == scenes.STYLE_GAS_GIANT:
        b = f.bodies[1]
        b["axis_x"][3] = 1.0
        b["surface_color"] = [0.2, 0.17, 0.13, 0.5]
        b["surface_params"] = [7.0, 1.0, 0.0, 0.0]
    return f


@pytest.mark.parametrize("style", STYLES)
def test_material_probe_matches_oracle(style):
    """Raw style albedo over the reference's 256x128 latitude-longitude grid.
    Both sides evaluate the noise hash bit for bit; elsewhere the GPU
    contracts multiply-adds, which moves a point across a cell or crater
    decision at a few points. Everywhere else the two agree to 1e-3, except on
    the edges of the icy lineae (see below)."""
    print(f"style {style}", flush=True)
    frame = style_frame(style)
    rays = scenes.material_rays(256, 128)
    gpu = PathTracer((4, 4), options=Options(milky_way=0.0)).probe(frame, PROBE_ALBEDO, rays)["normal"]
    ref = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f32").probe(frame, PROBE_ALBEDO, rays)["normal"]
    assert np.all(np.isfinite(gpu)) and np.all(gpu >= 0.0)
    assert not np.any(np.all(gpu == 0.0, axis=1))
    diff = np.abs(gpu.astype(np.float64) - ref.astype(np.float64)).max(axis=1)
    close = np.mean(diff < 1e-3)
    print(f"style {style}: points within 1e-3 {close:.4f}, median diff {np.median(diff):.2e}")
    if style == scenes.STYLE_ICY:
        # The icy lineae are 5e-4 to 4e-3 wide and their smoothstep edges
        # have slopes near 3e3, so a last-bit difference in a fracture's
        # distance moves a point on an edge by up to the lineae's contrast.
        # Off the edges the two agree; over the sphere the mean agrees.
        mean_diff = np.abs(gpu.astype(np.float64).mean(0) - ref.astype(np.float64).mean(0)).max()
        print(f"icy: within 0.1 {np.mean(diff < 0.1):.4f}, mean colour diff {mean_diff:.2e}")
        assert close >= 0.85, close
        assert np.mean(diff < 0.1) >= 0.97
        assert mean_diff < 2e-3
    else:
        assert close >= 0.98, close
    assert np.median(diff) < 1e-4


@pytest.mark.parametrize("styles", [(2, 3), (4, 5), (6, 2)], ids=["cratered-rocky", "icy-volcanic", "terran-cratered"])
def test_styled_moons_match_oracle_per_tile(styles):
    """Two styled moons rendered by both: per-tile 3-sigma test."""
    print(f"styled moons {styles}", flush=True)
    size = (64, 48)
    frame = scenes.styled_moons(*size, styles=styles)
    opts = Options(samples_per_frame=16, max_bounces=4, milky_way=0.0)
    gpu = PathTracer(size, options=opts)
    gpu.enable_moments(True)
    orc = oracle.OracleTracer(size, Options(samples_per_frame=16, max_bounces=4, milky_way=0.0), "f32")
    orc.enable_moments(True)
    for _ in range(2):
        gpu.encode(frame)
        orc.encode(frame)
    ga, gm, oa, om = gpu.read_accumulation(), gpu.read_moments(), orc.read_accumulation(), orc.read_moments()
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    assert np.mean(ga[..., :3].sum(-1) > 0) > 0.2
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    agree = parity.pixel_agreement(ga, oa)
    print(f"styled moons {styles}: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.99, agree


def test_style_codes_above_six_are_terran():
    """The WGSL switch sends every code above 5 to terran_albedo."""
    rays = scenes.material_rays(64, 32)
    frames = [scenes.material_frame(scenes.STYLE_TERRAN) for _ in range(2)]
    frames[1].bodies[1]["axis_x"][3] = 9.0
    t = PathTracer((4, 4), options=Options(milky_way=0.0))
    a, b = (t.probe(f, PROBE_ALBEDO, rays)["normal"] for f in frames)
    assert np.array_equal(a, b)


def test_styled_kernels_agree_across_integrators_and_batch():
    """Frames with solid-body styles take the kernels built with them, in
    every form: fused, wavefront and batched passes give bit-identical
    images (parity.assert_same_draws). A batch with one styled view runs all its views
    through the styled batch kernel."""
    from stargaze_amd.pathtracer import INTEGRATOR_FUSED, INTEGRATOR_WAVEFRONT, Batch

    size = (48, 36)
    frame = scenes.styled_moons(*size)
    images = []
    for integrator in (INTEGRATOR_FUSED, INTEGRATOR_WAVEFRONT):
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=4, milky_way=0.0))
        t.set_integrator(integrator)
        t.encode(frame)
        t.encode(frame)
        images.append(t.read_accumulation())
    parity.assert_same_draws(images[1], images[0])
    views = [PathTracer(size, options=Options(samples_per_frame=8, max_bounces=4, milky_way=0.0)) for _ in range(2)]
    plain = scenes.airless_triple(*size)
    batch = Batch(views)
    for _ in range(2):
        batch.encode([frame, plain])
    batch.synchronize()
    batched = [v.read_accumulation() for v in views]
    batch.close()
    parity.assert_same_draws(batched[0], images[0])
    ref = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=4, milky_way=0.0))
    ref.set_integrator(INTEGRATOR_FUSED)
    ref.encode(plain)
    ref.encode(plain)
    parity.assert_same_draws(batched[1], ref.read_accumulation())


def test_adjoint_rejects_solid_styles():
    """The adjoint does not differentiate the solid-body albedos (styles 2-6)."""
    size = (16, 12)
    frame = scenes.styled_moons(*size)
    t = PathTracer(size, options=Options(samples_per_frame=4, max_bounces=2, milky_way=0.0))
    t.enable_adjoint(True)
    t.encode(frame)
    with pytest.raises(NativeError, match="solid-body surface styles"):
        t.adjoint(frame, np.ones((size[1], size[0], 3), np.float32))
