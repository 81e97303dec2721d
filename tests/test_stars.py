"""Catalogue packing restated from src/stars.rs (CPU)."""
import math

import numpy as np

from stargaze_amd import stars


def test_every_direction_inside_a_disc_finds_its_star_in_its_cell():
    """The grid is conservative (src/stars.rs:150-151): a miss ray whose
    direction lies inside a star's disc, looked up the way the shader does
    (src/shaders/pathtrace.wgsl:285-296), must find that star, including
    discs across the longitude seam and over the poles."""
    rng = np.random.default_rng(3)
    catalogue = [stars.CatalogueStar(dir=(math.cos(p) * math.sin(t), math.sin(p) * math.sin(t), math.cos(t)),
                                     angular_radius=r)
                 for t, p, r in ((0.01, 0.3, 0.05), (math.pi - 0.02, 1.0, 0.04), (1.2, math.pi - 0.001, 0.03),
                                 (1.6, -math.pi + 0.002, 0.1), (0.7, 2.0, 1e-4), (2.3, -1.0, 0.02))]
    data, cells = stars.ray_catalogue(catalogue)
    assert cells.size == stars.GRID_W * stars.GRID_H + 1 + int(cells[stars.GRID_W * stars.GRID_H])
    for index, star in enumerate(catalogue):
        axis = np.asarray(star.dir) / np.linalg.norm(star.dir)
        assert abs(float(data["direction"][index][3]) - math.sin(star.angular_radius) ** 2) < 1e-9
        t1 = np.cross(axis, [0.0, 0.0, 1.0] if abs(axis[2]) < 0.9 else [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(axis, t1)
        for _ in range(400):
            rho = star.angular_radius * math.sqrt(rng.uniform()) * 0.999
            phi = rng.uniform(0.0, 2.0 * math.pi)
            d = math.cos(rho) * axis + math.sin(rho) * (math.cos(phi) * t1 + math.sin(phi) * t2)
            lon = math.atan2(d[1], d[0]) + math.pi
            theta = math.acos(max(-1.0, min(1.0, d[2])))
            x = min(int(lon * (256.0 / (2.0 * math.pi))), 255)
            y = min(int(theta * (128.0 / math.pi)), 127)
            cell = y * 256 + x
            listed = cells[32769 + cells[cell]:32769 + cells[cell + 1]]
            assert index in listed, (index, rho, phi, x, y)


def test_radiance_spreads_the_star_irradiance_over_its_disc():
    """src/stars.rs:132-139: radiance x pi r^2 = 6.4e-10 x 100 x bright / 1.62."""
    s = stars.CatalogueStar(dir=(0, 0, -1), color=(1.0, 0.5, 0.25), bright=3.0, angular_radius=0.02)
    rad = stars.radiance(s)
    flux = rad / np.array(s.color, np.float32) * math.pi * 0.02 ** 2
    assert np.allclose(flux, 6.4e-10 * 100.0 * 3.0 / 1.62, rtol=1e-5)
