"""The Milky Way panorama restated from src/sky.rs (CPU)."""
import math

import numpy as np

from stargaze_amd import sky


def test_galactic_haze_wraps_and_has_dust_contrast_without_pole_glow():
    """src/sky.rs:123-147: the longitude seam is continuous, the poles are
    dark, the central bulge is broader than the outer band, the core is warm
    and the outer clouds cool, and the map has lit and unlit texels."""
    for latitude in (-0.2, -0.05, 0.0, 0.05, 0.2):
        left, right = sky.color(-math.pi, latitude), sky.color(math.pi, latitude)
        assert np.linalg.norm(left - right) < 1e-6
    assert np.array_equal(sky.color(0.0, 1.5), np.zeros(3, np.float32))
    assert np.array_equal(sky.color(0.0, -1.5), np.zeros(3, np.float32))
    assert np.linalg.norm(sky.color(0.0, 0.35)) > 2.0 * np.linalg.norm(sky.color(math.pi, 0.35))
    warm, cool = sky.color(0.0, 0.25), sky.color(math.pi, 0.1)
    assert warm[0] > warm[2] and cool[2] > cool[0]
    m = sky.panorama()
    assert m.shape == (sky.HEIGHT, sky.WIDTH) and m.dtype == np.uint32
    assert np.any(m != 0) and np.any(m == 0)
    assert sky.panorama() is m, "the generated map is cached"


def test_panorama_packs_rgb10_of_the_colour_at_texel_centres():
    m = sky.panorama()
    for x, y in ((0, 256), (511, 250), (700, 300), (1023, 200)):
        lat = math.pi * (0.5 - (y + 0.5) / sky.HEIGHT)
        lon = 2 * math.pi * (0.5 - (x + 0.5) / sky.WIDTH)
        rgb = np.round(sky.color(lon, lat) * 1023.0).astype(np.uint32)
        assert m[y, x] == rgb[0] | (rgb[1] << 10) | (rgb[2] << 20)
