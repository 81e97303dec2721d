"""Ground themes 0-3 and 5 in the oracle (src/shaders/ground.wgsl:256-666):
every theme renders a finite, non-negative Halo frame, and each theme's
ground differs from the forest's."""
import numpy as np

import oracle
from stargaze_amd import scenes
from stargaze_amd.pathtracer import Options


def render(theme, size=(32, 18)):
    frame, data, cells = scenes.halo_parity(*size)
    frame.globals["ground_counts"][2] = theme
    t = oracle.OracleTracer(size, Options(samples_per_frame=2, max_bounces=2, milky_way=0.0), "f32")
    t.upload_terrain(data, cells)
    t.encode(frame)
    return t.read_accumulation()[..., :3].astype(np.float64)


def test_themes_render_finite_and_distinct():
    images = {theme: render(theme) for theme in (0, 1, 2, 3, 5)}
    for theme, image in images.items():
        assert np.all(np.isfinite(image)) and np.all(image >= 0.0), theme
        assert image.max() > 0.0, theme
    for theme in (1, 2, 3, 5):
        assert not np.allclose(images[theme], images[0]), theme
