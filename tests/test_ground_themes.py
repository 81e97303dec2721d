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


def test_ground_material_probe_on_the_oracle():
    """SG_PROBE_GROUND on the oracle: at ~1700 ground points per theme the
    albedo is finite and within [0, 1], the shading normal is a unit vector
    on the outer side of the surface, and every theme's floor varies from
    point to point; the points reach the terrain and the host sphere."""
    import ground_cases as gc
    from stargaze_amd.pathtracer import PROBE_GROUND

    for theme in (0, 1, 2, 3, 5):
        frame, data, cells = gc.theme_frame(theme)
        o64 = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f64")
        o64.upload_terrain(data, cells)
        rays = gc.material_rays(frame, o64)
        assert 0.1 < np.mean(rays["anchor"] == 0x80100000) < 0.9
        t = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f32")
        t.upload_terrain(data, cells)
        out = t.probe(frame, PROBE_GROUND, rays)
        a = out["extra"].astype(np.float64)
        n = out["normal"].astype(np.float64)
        assert np.all(np.isfinite(a)) and np.all((a >= 0.0) & (a <= 1.0)), theme
        assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5), theme
        up = frame.globals["ground_up"][:3].astype(np.float64)
        assert np.all(n @ up > 0.0), theme
        assert len(np.unique(a.round(4), axis=0)) > 0.5 * len(a), theme
