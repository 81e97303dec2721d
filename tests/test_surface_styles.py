"""Procedural surface styles 0-6 in the oracle: the reference's material probe
(src/surface_tests.rs:87-169) and closed forms of the cratered and rocky
albedos (src/shaders/pathtrace.wgsl:472-895).

Only the FP32 build is meaningful here: hash31 takes fract() of products of
coordinates offset by up to 512, so the noise is a function of FP32 rounding
by construction, and the FP64 build draws different (equally valid) noise."""
import numpy as np
import pytest

import oracle
from stargaze_amd import scenes
from stargaze_amd.pathtracer import PROBE_ALBEDO, Options

STYLES = tuple(range(7))  # plain, gas giant, cratered, rocky, icy, volcanic, terran


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


def luminance(c):
    return c @ np.array([0.2126, 0.7152, 0.0722])


def test_terran_sea_level_follows_the_ocean_fraction():
    """terran_albedo (wgsl:854-895): sea level is 0.5 + 0.3 (ocean - 0.5).
    With ocean = 1 (sea 0.65, above nearly all of the 7-octave fbm) and no
    ice or cloud, deep water is the ocean colour exactly: mix(., ocean, 1),
    then no land, ice or cloud. With ocean = 0 almost nothing is water."""
    rays = scenes.material_rays(128, 64)
    ocean = np.array([0.03, 0.09, 0.2])
    wet = probe(scenes.material_frame(scenes.STYLE_TERRAN, ocean=1.0, clouds=0.0, ice_caps=0.0), rays)
    dry = probe(scenes.material_frame(scenes.STYLE_TERRAN, ocean=0.0, clouds=0.0, ice_caps=0.0), rays)
    exact_wet = np.all(np.abs(wet - ocean) < 1e-7, axis=1).mean()
    exact_dry = np.all(np.abs(dry - ocean) < 1e-7, axis=1).mean()
    assert exact_wet > 0.9, exact_wet
    assert exact_dry < 0.05, exact_dry


def test_terran_cloud_cover_brightens():
    """Cloud mask threshold 0.62 - 0.22 cover - belts: more cover, more white."""
    rays = scenes.material_rays(128, 64)
    lum = [luminance(probe(scenes.material_frame(scenes.STYLE_TERRAN, clouds=c), rays)).mean()
           for c in (0.0, 0.4, 1.0)]
    assert lum[0] < lum[1] < lum[2], lum


def test_icy_mixes_are_channel_consistent():
    """icy_albedo mixes only the ice colour and tones of rust (0.8, 0.7 and
    1.0 times rust = secondary * (0.72, 0.62, 0.56)). With rust equal to the
    ice colour every output is the ice colour times one grey factor."""
    ice = np.array([0.8, 0.7, 0.6])
    secondary = ice / np.array([0.72, 0.62, 0.56])
    color = probe(scenes.material_frame(scenes.STYLE_ICY, albedo=tuple(ice), secondary=tuple(secondary)),
                  scenes.material_rays(128, 64))
    ratio = color / ice
    np.testing.assert_allclose(ratio[:, 1], ratio[:, 0], rtol=2e-5)
    np.testing.assert_allclose(ratio[:, 2], ratio[:, 0], rtol=2e-5)
    assert ratio.min() > 0.5 and ratio.max() < 1.2


def test_volcanic_paterae_follow_the_density():
    """volcanic_albedo: a cell holds a patera only if its hash is below
    0.6 x density, and a patera's pit blends 96% towards (0.035, 0.032,
    0.026). Density 0 leaves no such dark point; density 1 leaves some."""
    rays = scenes.material_rays(256, 128)
    bright = (0.85, 0.75, 0.3)
    none = luminance(probe(scenes.material_frame(scenes.STYLE_VOLCANIC, albedo=bright, craters=0.0), rays))
    dense = luminance(probe(scenes.material_frame(scenes.STYLE_VOLCANIC, albedo=bright, craters=1.0), rays))
    assert none.min() > 0.1, none.min()
    assert dense.min() < 0.06, dense.min()
    assert 0.001 < np.mean(dense < 0.06) < 0.5
