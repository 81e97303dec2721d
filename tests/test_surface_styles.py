"""Procedural surface styles 0-3 in the oracle: the reference's material probe
(src/surface_tests.rs:87-169) and closed forms of the cratered and rocky
albedos (src/shaders/pathtrace.wgsl:472-689).

Only the FP32 build is meaningful here: hash31 takes fract() of products of
coordinates offset by up to 512, so the noise is a function of FP32 rounding
by construction, and the FP64 build draws different (equally valid) noise."""
import numpy as np
import pytest

import oracle
from stargaze_amd import scenes
from stargaze_amd.pathtracer import PROBE_ALBEDO, Options

STYLES = (scenes.STYLE_PLAIN, scenes.STYLE_GAS_GIANT, scenes.STYLE_CRATERED, scenes.STYLE_ROCKY)


def probe(frame, rays, precision="f32"):
    out = oracle.OracleTracer((4, 4), Options(milky_way=0.0), precision).probe(frame, PROBE_ALBEDO, rays)
    return out["normal"].astype(np.float64)


def gas_giant_frame():
    f = scenes.material_frame(scenes.STYLE_PLAIN)
    b = f.bodies[1]
    b["axis_x"][3] = 1.0
    b["surface_color"] = [0.2, 0.17, 0.13, 0.5]
    b["surface_params"] = [7.0, 1.0, 0.0, 0.0]
    return f


def style_frame(style, **surface):
    return gas_giant_frame() if style == scenes.STYLE_GAS_GIANT else scenes.material_frame(style, **surface)


@pytest.mark.parametrize("style", STYLES)
def test_materials_are_finite_and_never_black(style):
    """src/surface_tests.rs:147-168: over a 256x128 latitude-longitude grid
    the raw material is finite, non-negative and never black."""
    color = probe(style_frame(style), scenes.material_rays(256, 128))
    assert np.all(np.isfinite(color))
    assert np.all(color >= 0.0)
    assert not np.any(np.all(color == 0.0, axis=1))


def test_rocky_polar_caps_cover_the_poles_only():
    """polar_caps (wgsl:650-656): with extent e the cap edge lies within
    1 - e +- 0.085 of |sin(latitude)|. The poles are the ice colour exactly,
    the equator band is ice-free; caps off leave no ice anywhere."""
    rays = scenes.material_rays(64, 32)
    z = rays["direction"][:, 2].astype(np.float64)
    ice = np.array([0.92, 0.92, 0.94])
    capped = probe(scenes.material_frame(scenes.STYLE_ROCKY, ice_caps=0.3), rays)
    bare = probe(scenes.material_frame(scenes.STYLE_ROCKY, ice_caps=0.0), rays)
    polar = np.abs(z) > 1.0 - 0.3 + 0.085 + 0.015
    equator = np.abs(z) < 1.0 - 0.3 - 0.085 - 0.015
    assert polar.any() and equator.any()
    np.testing.assert_allclose(capped[polar], np.broadcast_to(ice, capped[polar].shape), rtol=0, atol=1e-6)
    np.testing.assert_array_equal(capped[equator], bare[equator])
    assert not np.any(np.all(np.abs(bare - ice) < 1e-6, axis=1))


def test_crater_free_surfaces_are_grey_modulations():
    """With crater density 0 (craters() = 1) and the secondary albedo equal to
    the body's, the cratered albedo is the body albedo times the regolith
    grain, 0.9 + 0.2 fbm in [0.9, 1.1], the same factor in every channel."""
    albedo = (0.5, 0.4, 0.3)
    color = probe(scenes.material_frame(scenes.STYLE_CRATERED, albedo=albedo, craters=0.0, secondary=albedo),
                  scenes.material_rays(128, 64))
    ratio = color / np.array(albedo)
    np.testing.assert_allclose(ratio[:, 1], ratio[:, 0], rtol=1e-5)
    np.testing.assert_allclose(ratio[:, 2], ratio[:, 0], rtol=1e-5)
    assert ratio.min() >= 0.9 - 1e-5 and ratio.max() <= 1.1 + 1e-5
    assert ratio.max() - ratio.min() > 0.05, "the grain must vary"


def test_craters_darken_floors_and_brighten_rims_within_bounds():
    """The crater field multiplies the crater-free albedo. Per layer the floor
    factor is at least 1 - 0.18 c, so over the five layers (c = 1, 1, 0.85,
    0.7, 0.55) the product bound is 0.448. Relief adds rims, ejecta and rays,
    so some points brighten. A dense field darkens some floors."""
    rays = scenes.material_rays(256, 128)
    albedo = (0.5, 0.5, 0.5)
    with_craters = probe(scenes.material_frame(scenes.STYLE_CRATERED, albedo=albedo, craters=1.0, secondary=albedo),
                         rays)
    without = probe(scenes.material_frame(scenes.STYLE_CRATERED, albedo=albedo, craters=0.0, secondary=albedo),
                    rays)
    ratio = with_craters[:, 0] / without[:, 0]
    floor_bound = 0.82 * 0.82 * (1 - 0.18 * 0.85) * (1 - 0.18 * 0.7) * (1 - 0.18 * 0.55)
    assert ratio.min() >= floor_bound - 1e-4
    assert ratio.min() < 0.9, "crater floors must darken"
    assert ratio.max() > 1.05, "rims and rays must brighten"
    assert 0.02 < np.mean(ratio != 1.0) < 1.0
