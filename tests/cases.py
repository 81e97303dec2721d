"""The reference's transport checks, restated once and run on two tracers.

Each case follows a GPU test of the reference (file:line given per case) and
takes a ``make(dimensions, options)`` factory, so the same assertions pin the
CPU oracle (tests/test_oracle_golden.py) and the HIP library
(tests/test_gpu_transport.py). Closed forms come from the reference's own
assertions.
"""
from __future__ import annotations

import math

import numpy as np

from stargaze_amd import scenes
from stargaze_amd.pathtracer import Options

SUN_LIKE = dict(milky_way=0.0)


def opts(**kw) -> Options:
    """``test_options()`` (src/pathtracer_tests.rs:25-30): Milky Way off."""
    return Options(**{**SUN_LIKE, **kw})


def sample(tracer, frame, batches: int) -> np.ndarray:
    """``sample`` (src/pathtracer_tests.rs:184-197): encode, then read back."""
    for _ in range(batches):
        tracer.encode(frame)
    return tracer.read_accumulation().reshape(-1, 4)


# -- src/pathtracer_tests.rs:439-550 (gpu_transport) ----------------------------
def transport_emission(make):
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    f = scenes.frame()
    black = sample(t, f, 1)[0]
    assert np.array_equal(black, np.zeros(4, np.float32)), black
    f.bodies = scenes.bodies(scenes.body([0.0, 0.0, -1.0], 1e-6, [2.0, 3.0, 4.0], True))
    p = sample(t, f, 1)[0]
    assert np.array_equal(p, np.array([2, 3, 4, 0], np.float32)), p
    f.bodies["center"][0] = [0.0, 0.0, 0.0]
    f.bodies["radius"][0] = 1.0
    p = sample(t, f, 1)[0]
    assert np.array_equal(p, np.array([2, 3, 4, 0], np.float32)), p


def lit_scene():
    f = scenes.frame()
    f.bodies = scenes.bodies(
        scenes.body([0.0, 0.0, -3.0], 1.0, [0.8] * 3, False),
        scenes.body([0.0, 4.0, 2.0], 0.8, [10.0] * 3, True),
    )
    return f


def transport_direct(make):
    """Analytic irradiance of a uniform sphere: 0.8*10*0.8^2/32*sqrt(0.5) ±3%."""
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    full = sample(t, lit_scene(), 32)[0][0]
    expected = 0.8 * 10.0 * 0.8**2 / 32.0 * math.sqrt(0.5)
    assert abs(full - expected) < expected * 0.03, (full, expected)
    assert t.samples() == 2048
    return full


def transport_occluder_behind(make):
    """An occluder behind the emitter cannot shadow it.

    The reference compares two runs with continuing noise sequences at 1e-6;
    here both runs use the same sequence (fresh tracers), which makes the
    intended property (the occluder changes nothing) exact.
    """
    a = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    full = sample(a, lit_scene(), 32)[0][0]
    f = lit_scene()
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body([0.0, 8.0, 6.0], 1.0, [0.0] * 3, False))])
    b = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    behind = sample(b, f, 32)[0][0]
    assert abs(behind - full) < 1e-6, (behind, full)


def _with_occluder(center, radius):
    f = lit_scene()
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body(center, radius, [0.0] * 3, False))])
    return f


def transport_eclipses(make):
    """Umbra < 0.1%, penumbra 20-85%, annular 60-85% of the unshadowed value."""
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    full = sample(t, lit_scene(), 32)[0][0]
    umbra = sample(t, _with_occluder([0.0, 2.0, 0.0], 0.5), 32)[0][0]
    assert umbra < full * 0.001, (umbra, full)
    penumbra = sample(t, _with_occluder([0.4, 2.0, 0.0], 0.4), 32)[0][0]
    assert full * 0.2 < penumbra < full * 0.85, (penumbra, full)
    annular = sample(t, _with_occluder([0.0, 2.0, 0.0], 0.2), 32)[0][0]
    assert full * 0.6 < annular < full * 0.85, (annular, full)


def _binary():
    f = lit_scene()
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body([0.0, -4.0, 2.0], 0.8, [10.0] * 3, True))])
    return f


def transport_binary_and_mis(make):
    """Two suns add (2x ±5%); MIS with 8 bounces does not double count (±5%)."""
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    full = sample(t, lit_scene(), 32)[0][0]
    binary = sample(t, _binary(), 32)[0][0]
    assert abs(binary / full - 2.0) < 0.05, (binary, full)
    t.set_quality(64, 8, 1 << 24)
    bounced = sample(t, _binary(), 128)[0][0]
    assert abs(bounced / binary - 1.0) < 0.05, (bounced, binary)


def transport_indirect(make):
    """A white reflector lights a shadowed point; direct-only stays dark."""
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=8))
    f = scenes.frame()
    f.bodies = scenes.bodies(
        scenes.body([0.0, 0.0, -3.0], 1.0, [0.8] * 3, False),
        scenes.body([0.0, 4.0, 2.0], 0.8, [10.0] * 3, True),
        scenes.body([0.0, 2.0, 0.0], 0.5, [0.0] * 3, False),
        scenes.body([2.0, 0.0, -1.0], 1.0, [0.9] * 3, False),
    )
    indirect = sample(t, f, 128)[0][0]
    assert indirect > 1e-4, indirect
    t.set_quality(64, 1, 1 << 24)
    direct_only = sample(t, f, 32)[0][0]
    assert direct_only < 1e-6, direct_only


def transport_resize(make):
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    f = scenes.frame()
    sample(t, f, 1)
    t.resize((1, 1))
    assert t.samples() == 64
    t.resize((3, 2))
    assert t.samples() == 0
    f.globals["viewport"][:2] = [3.0, 2.0]
    assert sample(t, f, 1).shape == (6, 4)


# -- src/pathtracer_tests.rs:602-642 (gpu_distant_surface) ----------------------
def distant_surface(make, size=128):
    t = make((size, size), opts(samples_per_frame=16, max_bounces=1))
    f = scenes.frame()
    f.globals["viewport"] = [size, size, 1.0, 0.0]
    lo, hi = size * 100 // 256, size * 156 // 256
    for radius in (1e-4, 1e-6, 1e-7):
        f.globals["cam_forward"][3] = radius / 10.0 * 1.3
        f.bodies = scenes.bodies(
            scenes.body([0.0, 0.0, -10.0], radius, [0.8] * 3, False),
            scenes.body([5.0, 0.0, -5.0], 0.05, [10000.0] * 3, True),
        )
        pixels = sample(t, f, 4).reshape(size, size, 4)
        dark = int((pixels[lo:hi, lo:hi, 0] < 0.05).sum())
        assert dark == 0, (radius, dark)


# -- src/pathtracer_tests.rs:778-840 (gpu_surface_observer) ----------------------
def surface_observer(make, width=160, height=90):
    """Camera 2 m above an emissive host: no false sky hits, no ground misses."""
    t = make((width, height), opts(samples_per_frame=64, max_bounces=1))
    f = scenes.frame()
    fov = math.radians(90.0)
    f.globals["cam_forward"][3] = math.tan(fov * 0.5)
    f.globals["viewport"] = [width, height, 1.0, 0.0]
    radius = scenes.metres(6.4e6)
    # Zenith tilted 35 deg towards the camera's up, as at 35 deg altitude.
    alt = math.radians(35.0)
    zenith = np.array([0.0, math.cos(alt), math.sin(alt)])
    host = scenes.place(scenes.body([0, 0, 0], radius, [1.0] * 3, True), -zenith * (radius + scenes.metres(2.0)))
    f.bodies = scenes.bodies(host)
    pixels = sample(t, f, 1).reshape(height, width, 4)
    xs = (2.0 * (np.arange(width) + 0.5) / width - 1.0) * width / height
    ys = 1.0 - 2.0 * (np.arange(height) + 0.5) / height
    dirs = np.stack(np.broadcast_arrays(xs[None, :], ys[:, None], -np.ones((height, width))), -1)
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    altitude = dirs @ zenith
    value = pixels[..., 0]
    false_sky = int(((altitude > 0.02) & (value != 0.0)).sum())
    ground_miss = int(((altitude < -0.02) & (value != 1.0)).sum())
    assert false_sky == 0 and ground_miss == 0, (false_sky, ground_miss)


# -- src/ring_tests.rs:136-201 (gpu_rings_transport) ------------------------------
def rings_transport(make):
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    f = scenes.frame()
    ring = scenes.ringed([1.8, 0.0, -5.0], 1.0, [0.0, 0.0, 1.0], 0.7)
    ring["ring_color"] = [0.0] * 4
    f.bodies = scenes.bodies(ring, scenes.body([0.0, 0.0, -9.0], 0.5, [1.0] * 3, True))
    for pole, expected in (([0.0, 0.0, 1.0], math.exp(-0.7)), ([0.0, 0.8, 0.6], math.exp(-0.7 / 0.6))):
        f.bodies["ring_plane"][0][:3] = pole
        p = sample(t, f, 256)[0]
        assert abs(p[0] - expected) < 0.02, (p, expected)
    f.bodies = scenes.bodies(scenes.ringed([1.8, 0.0, -5.0], 1.0, [0.0, 0.0, 1.0], 0.7), f.bodies[1])
    for star_z, same_side in ((15.0, True), (-25.0, False)):
        f.bodies[1] = scenes.body([10.0, 0.0, star_z], 0.05, [100000.0] * 3, True)
        mu = 20.0 / math.sqrt(500.0)
        if same_side:
            factor = (1.0 - math.exp(-0.7 * (1.0 + 1.0 / mu))) / (1.0 + mu)
        else:
            factor = (math.exp(-0.7 / mu) - math.exp(-0.7)) / (mu - 1.0)
        s2 = 0.05**2 / 500.0
        solid_angle = 2.0 * math.pi * s2 / (1.0 + math.sqrt(1.0 - s2))
        expected = 0.8 * factor / (4.0 * math.pi) * mu * 100000.0 * solid_angle
        p = sample(t, f, 32)[0]
        assert abs(p[0] / expected - 1.0) < 0.02, (p, expected, same_side)
        blocked = scenes.frame()
        blocked.bodies = np.concatenate(
            [f.bodies, scenes.bodies(scenes.body([5.0, 0.0, (star_z - 5.0) * 0.5], 0.5, [0.0] * 3, False))]
        )
        dark = sample(t, blocked, 8)[0]
        assert dark[0] < p[0] * 0.001, (dark, p)


# -- src/ring_tests.rs:414-452 (gpu_rings_null_crossings_preserve_mis) ----------
def rings_null_mis(make, size=32):
    t = make((size, size), opts(samples_per_frame=64, max_bounces=1))
    f = scenes.frame()
    f.globals["viewport"][:2] = [size, size]
    ring = scenes.ringed([0.0, 4.0, -2.0], 0.1, [0.0, 0.0, 1.0], 0.4)
    ring["ring_params"][0] = 100.0
    ring["ring_color"] = [0.0] * 4
    f.bodies = scenes.bodies(
        scenes.body([0.0, 0.0, -5.0], 1.0, [0.8] * 3, False), scenes.body([0.0, 4.0, 2.0], 3.8, [1.0] * 3, True), ring
    )
    direct = sample(t, f, 4)[..., 0].astype(np.float64).mean()
    t.set_quality(64, 2, 1 << 24)
    mis = sample(t, f, 4)[..., 0].astype(np.float64).mean()
    assert abs(mis / direct - 1.0) < 0.01, (mis, direct)


# -- src/pathtracer_tests.rs:842-934 (gpu_atmosphere, extinction and eclipse) ----
def earthlike():
    """``Atmosphere::earthlike`` (src/sim.rs:295-306), in AU units."""
    au_m = scenes.METRES_PER_AU
    return dict(scale_height=8500.0 / au_m, thickness=80000.0 / au_m,
                rayleigh=[5.8e-6 * au_m, 13.5e-6 * au_m, 33.1e-6 * au_m], mie=21.0e-6 * au_m, mie_g=0.76)


def air_frame(a):
    f = scenes.frame()
    radius = np.float32(6371.0 / scenes.AU_KM)
    center = [0.0, 0.0, float(radius + np.float32(0.002 / scenes.AU_KM))]
    f.bodies = scenes.bodies(scenes.body(center, radius, [0.0] * 3, False))
    f.globals["atmo_center"] = [center[0], center[1], center[2], radius]
    f.globals["atmo_rayleigh"] = [*a["rayleigh"], a["mie"]]
    f.globals["atmo_params"] = [a["mie_g"], a["scale_height"], a["thickness"], 1.0]
    return f


def atmosphere_zenith_and_eclipse(make):
    """Zenith extinction of a catalogue star: exp(-beta * midpoint column), 0.1%;
    an eclipsing body leaves the air exactly dark."""
    a = earthlike()
    # One catalogue star straight up (CatalogueStar dir (0,0,-1), radius 0.01).
    radiance = 6.4e-10 * 100.0 / (math.pi * 0.01 * 0.01)
    cat = np.zeros(1, [("direction", "<f4", (4,)), ("radiance", "<f4", (4,))])
    cat["direction"][0] = [0.0, 0.0, -1.0, math.sin(0.01) ** 2]
    cat["radiance"][0] = [radiance, radiance, radiance, 0.0]
    cells = sky_cells_for([(np.array([0.0, 0.0, -1.0]), 0.01)])
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1), catalogue=cat, sky_cells=cells)
    f = air_frame(a)
    att = sample(t, f, 1)[0]
    x_km = 0.002 / scenes.AU_KM
    column = a["scale_height"] * (math.exp(-x_km / a["scale_height"]) - math.exp(-a["thickness"] / a["scale_height"]))
    x = (a["thickness"] - x_km) / (8.0 * a["scale_height"])
    midpoint_column = column * x / (2.0 * math.sinh(x * 0.5))
    for ch in range(3):
        expected = math.exp(-(a["rayleigh"][ch] + a["mie"]) * midpoint_column) * radiance
        assert abs(att[ch] - expected) < 0.001 * expected, (ch, att[ch], expected)
    f.globals["atmo_params"][3] = 0.0
    un = sample(t, f, 1)[0]
    for ch in range(3):
        assert abs(un[ch] - radiance) <= 1e-5 * radiance, (un, radiance)
    assert un[3] == 0.0
    # Daylight sky is blue; an occluder shadows the air completely.
    t2 = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    f = air_frame(a)
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body([1.0, 0.0, -1.0], 0.005, [100000.0] * 3, True))])
    day = sample(t2, f, 4)[0]
    assert np.all(np.isfinite(day[:3])) and np.all(day[:3] > 0.01), day
    assert day[2] > day[0], day
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body([0.5, 0.0, -0.5], 0.05, [0.0] * 3, False))])
    eclipsed = sample(t2, f, 4)[0]
    assert np.array_equal(eclipsed, np.zeros(4, np.float32)), eclipsed


def atmosphere_split_heights_zenith(make):
    """Extension of the zenith pin to the 4K workload's air model: 16 view
    cells and a Mie density with its own 1.2 km scale height. Each species'
    column is then its own midpoint sum, in closed form
    H (e^{-h0/H} - e^{-T/H}) * x / (2 sinh(x/2)) with x = dt / H, and the
    transmittance is exp(-(beta_r * column_r + beta_m * column_m)), 0.1%. A
    Mie height equal to the Rayleigh one gives the one-height image (1e-5)."""
    a = earthlike()
    au = scenes.METRES_PER_AU
    radiance = 6.4e-10 * 100.0 / (math.pi * 0.01 * 0.01)
    cat = np.zeros(1, [("direction", "<f4", (4,)), ("radiance", "<f4", (4,))])
    cat["direction"][0] = [0.0, 0.0, -1.0, math.sin(0.01) ** 2]
    cat["radiance"][0] = [radiance, radiance, radiance, 0.0]
    cells = sky_cells_for([(np.array([0.0, 0.0, -1.0]), 0.01)])
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1), catalogue=cat, sky_cells=cells)
    f = air_frame(a)
    mie_h = 1200.0 / au
    t.set_air_model(16, 8, mie_h)
    att = sample(t, f, 1)[0]
    h0 = 0.002 / scenes.AU_KM
    dt = (a["thickness"] - h0) / 16.0

    def midpoint_column(h):
        x = dt / h
        return h * (math.exp(-h0 / h) - math.exp(-a["thickness"] / h)) * x / (2.0 * math.sinh(0.5 * x))

    col_r, col_m = midpoint_column(a["scale_height"]), midpoint_column(mie_h)
    for ch in range(3):
        expected = math.exp(-(a["rayleigh"][ch] * col_r + a["mie"] * col_m)) * radiance
        assert abs(att[ch] - expected) < 0.001 * expected, (ch, att[ch], expected)
    # Equal heights: the split model is the one-height model.
    t.set_air_model(16, 8, float(np.float32(a["scale_height"])))
    same = sample(t, f, 1)[0]
    t.set_air_model(16, 8, 0.0)
    one = sample(t, f, 1)[0]
    assert np.allclose(same[:3], one[:3], rtol=1e-5, atol=0.0), (same, one)
    t.set_air_model()


# -- src/ring_tests.rs:55-134 (gpu_rings_geometry) -------------------------------
def _camera_probe(direction, target):
    from stargaze_amd.layouts import MISS, PROBE_RAY

    ray = np.zeros(1, PROBE_RAY)
    ray["anchor"] = MISS
    ray["direction"] = direction
    ray["target"] = target
    return ray


def rings_geometry_probe(make):
    """The production ring helpers on one camera ray: a ring 10 AU away, its
    pole tilted to the ray, is hit at t = 10 (1e-5) and radius 1.8 planet
    radii (1e-4) with transmission and star visibility exp(-0.7 / 0.6)
    (1e-5); the inner hole, the outside and an edge-on ring miss; a star in
    front of the ring is not attenuated (1e-6)."""
    t = make((1, 1), opts())
    r = 0.0004
    f = scenes.frame()
    f.bodies = scenes.bodies(scenes.ringed([1.8 * r, 0.0, -10.0], r, [0.0, 0.8, 0.6], 0.7),
                             scenes.body([0.0, 0.0, -12.0], 0.1, [1.0] * 3, True))
    ring = t.probe(f, 2, _camera_probe([0.0, 0.0, -1.0], 0))[0]
    vis = t.probe(f, 3, _camera_probe([0.0, 0.0, -1.0], 1))[0]
    expected = math.exp(-0.7 / 0.6)
    radius = float(np.linalg.norm(np.asarray(ring["extra"], np.float64))) / r
    assert int(ring["index"]) == 0
    assert abs(float(ring["distance"]) - 10.0) < 1e-5 and abs(radius - 1.8) < 1e-4, (ring, radius)
    assert abs(float(ring["normal"][0]) - expected) < 1e-5, ring
    assert abs(float(vis["distance"]) - expected) < 1e-5, vis
    for offset, pole in ((0.5, [0.0, 0.0, 1.0]), (2.6, [0.0, 0.0, 1.0]), (1.8, [0.0, 1.0, 0.0])):
        f.bodies[0] = scenes.ringed([offset * r, 0.0, -10.0], r, pole, 0.7)
        miss = t.probe(f, 2, _camera_probe([0.0, 0.0, -1.0], 0))[0]
        assert int(miss["index"]) == 0xFFFFFFFF, (offset, pole, miss)
    f.bodies[0] = scenes.ringed([1.8 * r, 0.0, -10.0], r, [0.0, 0.0, 1.0], 0.7)
    f.bodies["center"][1] = [0.0, 0.0, -9.0]
    assert abs(float(t.probe(f, 3, _camera_probe([0.0, 0.0, -1.0], 1))[0]["distance"]) - 1.0) < 1e-6


# -- src/pathtracer_tests.rs:997-1094 (gpu_reflected_atmosphere) ----------------
def planetshine_reflected_air(make):
    """Planetshine in the air: a sunlit moon beside the zenith lights the air
    (> 1e-4); its glow scales with the moon's albedo tint (2%), shrinks with
    its apparent disc (0.15-0.35 at half the radius), follows its phase
    (quarter 0.1-0.6, new < 2%), and vanishes when the moon is eclipsed, when
    a body hides the moon from the air, and without air (< 1e-6)."""
    a = earthlike()
    f = air_frame(a)
    center = list(f.bodies["center"][0])
    radius = float(f.bodies["radius"][0])
    f.bodies = scenes.bodies(scenes.body(center, radius, [0.0] * 3, False),
                             scenes.body([0.0004, 0.0, -0.002], 0.00015, [0.8] * 3, False),
                             scenes.body([0.3, 0.0, 1.0], 0.005, [100000.0] * 3, True))

    def glow(frame):
        # Common random numbers: both renders start a fresh tracer at the
        # same noise sequence, so the difference is the moon's light alone
        # and not the sky's Monte Carlo noise (the cone-sampled sun moves
        # the Mie phase by about 1% per sample in daylight).
        color = frame.bodies["color"][1].copy()
        frame.bodies["color"][1] = [0.0] * 3
        background = sample(make((1, 1), opts(samples_per_frame=64, max_bounces=1)), frame, 32)[0]
        frame.bodies["color"][1] = color
        lit = sample(make((1, 1), opts(samples_per_frame=64, max_bounces=1)), frame, 32)[0]
        return lit[:3].astype(np.float64) - background[:3].astype(np.float64)

    full = glow(f)
    assert np.all(np.isfinite(full)) and np.all(full > 1e-4), full
    f.bodies["color"][1] = [0.4, 0.2, 0.1]
    tinted = glow(f)
    for c in range(3):
        expected = full[c] * (0.5, 0.25, 0.125)[c]
        assert abs(tinted[c] - expected) < 0.02 * expected, (tinted, full)
    f.bodies["color"][1] = [0.8] * 3
    f.bodies["radius"][1] *= 0.5
    small = glow(f)
    assert full[0] * 0.15 < small[0] < full[0] * 0.35, (small, full)
    f.bodies["radius"][1] *= 2.0
    f.bodies["center"][2] = [1.0, 0.0, 0.0]
    quarter = glow(f)
    f.bodies["center"][2] = [-0.3, 0.0, -1.0]
    new = glow(f)
    assert full[0] * 0.1 < quarter[0] < full[0] * 0.6, (quarter, full)
    assert abs(new[0]) < full[0] * 0.02, (new, full)
    f.bodies["center"][2] = [0.3, 0.0, 1.0]
    base = f.bodies.copy()
    f.bodies = np.concatenate([base, scenes.bodies(scenes.body([0.1502, 0.0, 0.499], 0.01, [0.0] * 3, False))])
    eclipsed = glow(f)
    assert np.all(np.abs(eclipsed) < 1e-6), eclipsed
    f.bodies = np.concatenate([base, scenes.bodies(scenes.body([0.0002, 0.0, -0.001], 0.00012, [0.0] * 3, False))])
    hidden = glow(f)
    assert np.all(np.abs(hidden) < 1e-6), hidden
    f.bodies = base
    f.globals["atmo_params"][3] = 0.0
    airless = glow(f)
    assert np.all(np.abs(airless) < 1e-6), airless


# -- src/pathtracer_tests.rs:1204-1266 (gpu_star_dots) ---------------------------
def star_dots(make):
    """An off-axis catalogue star of 0.1 rad at (0.8, 0, -1) in a 90 deg
    128^2 view is capped to 3 pixels around its image point (115.2, 64):
    every lit pixel lies within 3 + 1/sqrt(2) px and more than 20 are lit.
    A bright catalogue star behind the camera lights no dark ground, at
    either zoom."""
    from stargaze_amd.stars import CatalogueStar, ray_catalogue

    star = CatalogueStar(dir=(0.8, 0.0, -1.0), color=(1.0, 1.0, 1.0), bright=1.0, angular_radius=0.1)
    cat, cells = ray_catalogue([star])
    t = make((128, 128), opts(samples_per_frame=64, max_bounces=2, exposure=1.0), catalogue=cat, sky_cells=cells)
    f = scenes.frame()
    f.globals["cam_forward"][3] = 1.0
    f.globals["viewport"] = [128.0, 128.0, 1.0, 0.0]
    pixels = sample(t, f, 1)
    lit = 0
    for i, p in enumerate(pixels):
        if p[0] > 0.0:
            dx = (i % 128) + 0.5 - 115.2
            dy = (i // 128) + 0.5 - 64.0
            assert math.hypot(dx, dy) < 3.0 + math.sqrt(0.5), (i, p)
            lit += 1
    assert lit > 20, lit
    behind = CatalogueStar(dir=(0.0, 0.0, 1.0), color=(1.0, 1.0, 1.0), bright=100.0, angular_radius=0.1)
    cat, cells = ray_catalogue([behind])
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=2, exposure=1.0), catalogue=cat, sky_cells=cells)
    f = scenes.frame()
    f.bodies = scenes.bodies(scenes.body([0.0, 0.0, -3.0], 1.0, [0.8] * 3, False))
    before = sample(t, f, 128)[0][0]
    f.globals["cam_forward"][3] *= 0.01
    after = sample(t, f, 128)[0][0]
    assert before == 0.0 and after == 0.0, (before, after)


# -- src/sky_tests.rs:23-103 (gpu_milky_way_is_camera_only_...) ------------------
def galactic_frame():
    """A camera looking at the galactic centre (src/sky_tests.rs:3-11): the
    view direction is the panorama's zero-longitude axis."""
    f = scenes.frame()
    pole = np.array([0.35, 0.82, 0.45])
    pole /= np.linalg.norm(pole)
    sign = 1.0 if pole[2] >= 0 else -1.0
    a = -1.0 / (sign + pole[2])
    direction = np.array([pole[0] * pole[1] * a, sign + pole[1] * pole[1] * a, -pole[1]])
    f.globals["cam_forward"][:3] = direction
    f.globals["cam_right"][:3] = np.cross(direction, pole)
    f.globals["cam_up"][:3] = pole
    return f


def milky_way_camera_only(make):
    """The Milky Way is seen by camera rays only: the galactic core gives
    0.01..1 x the backdrop gain, stays warm (r > 1.2 g), and its luminance is
    the backdrop channel w (1e-7); a diffuse body under it stays exactly
    black; an emissive body in front occludes it (its own radiance, w = 0);
    after a resize the panorama is unchanged (1%). With automatic exposure
    the backdrop does not drive the meter (1e-4) and the night exposure is
    capped in [200, 300] (where the tracer meters)."""
    from stargaze_amd.sky import panorama

    o = opts(samples_per_frame=64, auto_exposure=True)
    o.milky_way = 4.0e-4  # Options::default()
    t = make((1, 1), o, milky_way=panorama())
    f = galactic_frame()
    night = sample(t, f, 64)[0]
    gain = 4.0e-4
    assert 0.01 * gain < night[0] <= gain, night
    lum = night[0] * 0.2126 + night[1] * 0.7152 + night[2] * 0.0722
    assert abs(lum - night[3]) < 1e-7, night
    assert night[0] > night[1] * 1.2, night
    meters = hasattr(t, "read_exposure")
    if meters:
        with_backdrop = float(t.read_exposure()[0]["value"])
        o2 = opts(samples_per_frame=64, auto_exposure=True)
        plain = make((1, 1), o2, milky_way=panorama())
        sample(plain, f, 64)
        without = float(plain.read_exposure()[0]["value"])
        assert abs(with_backdrop - without) <= 1e-4 * without, (with_backdrop, without)
    f.bodies = scenes.bodies(scenes.body([0.0, 0.0, -3.0], 1.0, [0.9] * 3, False))
    f.globals["cam_forward"][:3] = [0.0, 0.0, -1.0]
    f.globals["cam_right"][:3] = [1.0, 0.0, 0.0]
    f.globals["cam_up"][:3] = [0.0, 1.0, 0.0]
    dark = sample(t, f, 16)[0]
    assert np.array_equal(dark, np.zeros(4, np.float32)), dark
    f.bodies["emissive"][0] = 1.0
    f.bodies["color"][0] = [1e-5] * 3
    dim = sample(t, f, 64)[0]
    assert np.all(np.abs(dim[:3] - 1e-5) < 1e-10) and dim[3] == 0.0, dim
    if meters:
        capped = float(t.read_exposure()[0]["value"])
        assert 200.0 <= capped <= 300.0, capped
    t.resize((2, 2))
    f = galactic_frame()
    f.globals["viewport"][:2] = [2.0, 2.0]
    resized = sample(t, f, 1)
    assert np.all(np.abs(resized[:, 0] - night[0]) < night[0] * 0.01), (resized, night)


# -- src/ground_tests.rs:1053-1153 (gpu_shadow_queries_match_closest_hits) -------
def _reference_visibility(t, f, rays, light):
    """closest_hit must be the light; then the product of the ring
    transmissions in front of it (the reference's reference_visibility)."""
    closest = t.probe(f, 1, rays)
    out = np.where(closest["index"].astype(np.uint32) == light, 1.0, 0.0)
    for i in range(len(f.bodies)):
        r = rays.copy()
        r["target"] = i
        ring = t.probe(f, 2, r)
        before = (ring["index"].astype(np.uint32) != 0xFFFFFFFF) & (ring["distance"] < closest["distance"])
        out = np.where(before, out * ring["normal"][:, 0], out)
    return out


def shadow_queries_match_closest_hits(make):
    """The any-hit shadow query (light_visibility) equals the closest-hit
    reference bit for bit: a ringed planet, two lights at one place (the
    second is hidden by the tie rule) and a body behind them; then a scene
    with a landscape (terrain, props, host sphere) from the surface, 2 km up,
    1 km under ground and three host radii out, toward each light and in
    random directions."""
    from stargaze_amd.layouts import MISS, PROBE_RAY

    rng = np.random.default_rng(7)
    ring_scene = scenes.frame()
    planet = scenes.ringed([1.8, 0.0, -5.0], 1.0, [0.0, 0.8, 0.6], 0.7)
    planet["ring_params"][0] = 2.4
    ring_scene.bodies = scenes.bodies(planet, scenes.body([0.0, 0.0, -10.0], 0.5, [1.0] * 3, True),
                                      scenes.body([0.0, 0.0, -10.0], 0.5, [1.0] * 3, True),
                                      scenes.body([0.0, 0.0, -20.0], 1.0, [0.5] * 3, False))
    halo, data, cells = scenes.halo_parity(64, 16)
    cases_ = [(ring_scene, None), (halo, (data, cells))]
    for f, terrain in cases_:
        t = make((4, 4), opts())
        if terrain is not None:
            t.upload_terrain(*terrain)
        g = f.globals
        origins = [np.zeros(3)]
        if g["ground_north"][3] != 0.0:
            east, up, north = (g[k][:3].astype(np.float64) for k in ("ground_east", "ground_up", "ground_north"))
            origins = [up * h / scenes.METRES_PER_AU for h in (0.0, 2000.0, -1000.0, 3.0 * float(g["ground_up"][3]))]
        centres = f.bodies["center"].astype(np.float64) + f.bodies["center_low"][:, :3].astype(np.float64)
        seen = []
        for light in range(len(f.bodies)):
            if f.bodies["emissive"][light] < 0.5:
                continue
            rays = np.zeros(64 * len(origins), PROBE_RAY)
            k = 0
            for o in origins:
                for j in range(64):
                    if j % 2 == 0:
                        d = centres[light] - o
                        d /= np.linalg.norm(d)
                        d = d + rng.normal(size=3) * float(f.bodies["radius"][light]) / np.linalg.norm(centres[light] - o)
                    else:
                        d = rng.uniform(-0.5, 0.5, 3)
                    rays[k]["offset"] = o
                    rays[k]["anchor"] = MISS
                    rays[k]["direction"] = d / np.linalg.norm(d)
                    rays[k]["target"] = light
                    k += 1
            vis = t.probe(f, 3, rays)["distance"].astype(np.float64)
            ref = _reference_visibility(t, f, rays, light)
            assert np.array_equal(vis, ref), (light, np.nonzero(vis != ref)[0][:8], vis[vis != ref][:8],
                                              ref[vis != ref][:8])
            seen.append(vis)
        seen = np.concatenate(seen)
        assert 0 < np.count_nonzero(seen) < len(seen)
        if terrain is None:
            assert np.any((seen > 0) & (seen < 1)), "some rays must cross the ring"


# -- src/ground_tests.rs:191-272 (gpu_local_ground_shadows) ---------------------
def local_frame(radius_m):
    """``local_frame`` (src/ground_tests.rs:191-206): the camera 2 m above a
    host of radius_m, the ground plane tilted to the view, a 0.5 m sun 8 m up
    and 4 m south."""
    metre = 1.0 / scenes.METRES_PER_AU
    f = scenes.frame()
    up = np.array([0.0, math.sqrt(0.8), math.sqrt(0.2)], np.float32)
    north = np.array([0.0, -up[2], up[1]], np.float32)
    f.globals["ground_east"] = [1.0, 0.0, 0.0, 2.0]
    f.globals["ground_up"] = [*up, radius_m]
    f.globals["ground_north"] = [*north, 1.0]
    center = -up.astype(np.float64) * (radius_m + 2.0) * metre
    sun = (up.astype(np.float64) * 8.0 - north.astype(np.float64) * 4.0) * metre
    f.bodies = scenes.bodies(scenes.body(center, radius_m * metre, [0.3] * 3, False),
                             scenes.body(sun, 0.5 * metre, [100.0] * 3, True))
    return f


def _one_prop(f, center, extent):
    from stargaze_amd.layouts import PRIMITIVE

    p = np.zeros(1, PRIMITIVE)
    p["center"] = center
    p["extent"] = extent
    f.ground = p
    f.globals["ground_counts"][0] = 1


def local_ground_shadows(make):
    """A box, an ellipsoid, a rotated box and a blocker 2 cm above the ground
    shadow the sun (< 0.1%, < 1% for the blocker) on two host radii."""
    t = make((1, 1), opts(samples_per_frame=64, max_bounces=1))
    for radius in (6_371_000.0, 1_500_000.0):
        f = local_frame(radius)
        full = sample(t, f, 8)[0][0]
        assert 0.01 < full < np.inf, full
        _one_prop(f, [0.0, 4.0, -4.0, 0.0], [0.8, 0.8, 0.8, 0.0])
        box = sample(t, f, 8)[0][0]
        assert box < full * 0.001, (box, full)
        f.ground["center"][0][3] = 1.0
        ellipsoid = sample(t, f, 8)[0][0]
        assert ellipsoid < full * 0.001, (ellipsoid, full)
        _one_prop(f, [0.0, 0.02, -4.0, 0.0], [0.2, 0.004, 0.005, 0.0])
        small = sample(t, f, 8)[0][0]
        assert small < full * 0.01, (small, full)
        _one_prop(f, [0.0, 4.0, -4.0, 0.0], [0.8, 0.8, 0.8, 0.4])
        rotated = sample(t, f, 8)[0][0]
        assert rotated < full * 0.001, (rotated, full)


# -- src/ground_tests.rs:630-677 (gpu_distant_hills_and_rocks_occlude_the_sky) -
def distant_props_occlude_the_sky(make):
    """Ellipsoid, rock, conifer and log primitives 450 m away and 35-60 m in
    size block a sun behind them (< 1e-6) and let it through when moved
    100 m aside (> 0.99), on two host radii."""
    t = make((1, 1), opts(samples_per_frame=1, max_bounces=1))
    metre = 1.0 / scenes.METRES_PER_AU
    alt = math.radians(3.0)
    up = np.array([0.0, math.cos(alt), -math.sin(alt)])
    north = np.array([0.0, -math.sin(alt), -math.cos(alt)])
    for radius in (1_500_000.0, 6_371_000.0):
        f = scenes.frame()
        f.globals["ground_east"] = [1.0, 0.0, 0.0, 2.0]
        f.globals["ground_up"] = [*up, radius]
        f.globals["ground_north"] = [*north, 1.0]
        f.bodies = scenes.bodies(scenes.body(-up * (radius + 2.0) * metre, radius * metre, [0.0] * 3, False),
                                 scenes.body([0.0, 0.0, -1.0], 0.1, [1.0] * 3, True))
        assert sample(t, f, 1)[0][0] > 0.99
        for kind in (1.0, 2.0, 3.0, 5.0):
            _one_prop(f, [0.0, 25.0, 450.0, kind], [35.0, 30.0, 60.0, 0.2])
            assert sample(t, f, 1)[0][0] < 1e-6, (radius, kind)
            f.ground["center"][0][0] = 100.0
            assert sample(t, f, 1)[0][0] > 0.99, (radius, kind)


def sky_cells_for(stars):
    """``ray_catalogue`` cell lists (src/stars.rs:152-205) for a few stars."""
    W, H = 256, 128
    bins = [[] for _ in range(W * H)]
    for index, (d, radius) in enumerate(stars):
        d = d / np.linalg.norm(d)
        theta = math.acos(max(-1.0, min(1.0, d[2])))
        phi = math.atan2(d[1], d[0]) + math.pi
        y0 = min(int(max(theta - radius, 0.0) / math.pi * H), H - 1)
        y1 = min(int(min(theta + radius, math.pi) / math.pi * H), H - 1)
        if theta <= radius or theta + radius >= math.pi:
            lon = math.pi
        else:
            lon = math.asin(max(0.0, min(1.0, math.sin(radius) / math.sin(theta))))
        for x in range(W):
            cell_phi = (x + 0.5) * 2 * math.pi / W
            sep = abs(((cell_phi - phi + math.pi) % (2 * math.pi)) - math.pi)
            if sep <= lon + math.pi / W + 1e-6:
                for y in range(y0, y1 + 1):
                    bins[y * W + x].append(index)
    offsets, indices = [], []
    for b in bins:
        offsets.append(len(indices))
        indices.extend(b)
    offsets.append(len(indices))
    return np.array(offsets + indices, np.uint32)


ALL = [
    transport_emission,
    transport_direct,
    transport_occluder_behind,
    transport_eclipses,
    transport_binary_and_mis,
    transport_indirect,
    transport_resize,
    distant_surface,
    surface_observer,
    rings_transport,
    rings_null_mis,
    atmosphere_zenith_and_eclipse,
    atmosphere_split_heights_zenith,
    rings_geometry_probe,
    planetshine_reflected_air,
    star_dots,
    milky_way_camera_only,
    shadow_queries_match_closest_hits,
    local_ground_shadows,
    distant_props_occlude_the_sky,
]


# -- src/motion_noise_tests.rs:40-85,172-200,326-405 (gpu_eclipse_energy_and_sample_batching)
def eclipsed_planetshine(size, step):
    """A telescope view (5 deg half-height) from 2 m above an Earth-sized host
    with Earth-like air, a small moon low in the view, a sun behind the
    observer and a large dark body moving across it by half a kilometre
    per step (src/motion_noise_tests.rs:40-85)."""
    a = earthlike()
    km = 1.0 / scenes.AU_KM
    radius = np.float32(6371.0 * km)
    center = [0.0, 0.0, float(radius + np.float32(0.002 * km))]
    f = scenes.frame()
    f.globals["viewport"] = [size[0], size[1], 1.0, size[0] / size[1]]
    f.globals["cam_forward"][3] = math.tan(math.radians(5.0))
    f.globals["ground_up"] = [0.0, 0.0, -1.0, 6371000.0]
    f.globals["atmo_center"] = [0.0, 0.0, center[2], radius]
    f.globals["atmo_rayleigh"] = [*a["rayleigh"], a["mie"]]
    f.globals["atmo_params"] = [a["mie_g"], a["scale_height"], a["thickness"], 1.0]
    d = 100_000.0 * km
    f.bodies = scenes.bodies(
        scenes.body(center, radius, [0.0] * 3, False),
        scenes.body([0.0004 + step * 0.00001, 0.0, -0.002], 0.00004, [0.8, 0.6, 0.4], False),
        scenes.body([-0.8, 0.0, -0.6], 0.005, [1e7] * 3, True),
        scenes.body([-0.8 * d + step * 0.5 * km, 0.0, -0.6 * d], 1100.0 * km, [0.0] * 3, False),
    )
    f.scene_time = float(step)
    return f


def eclipsed_sun(size, step):
    """The same host seen at a 5 deg elevation (4 deg half-height) with the sun
    30 deg above the horizon behind a 513 km body 1e5 km away that moves
    across it (src/motion_noise_tests.rs:172-200)."""
    f = eclipsed_planetshine(size, step)
    km = 1.0 / scenes.AU_KM
    e = math.radians(5.0)
    up = np.array([0.0, math.cos(e), -math.sin(e)])
    north = np.array([0.0, math.sin(e), math.cos(e)])
    center = -up * ((6371.0 + 0.002) * km)
    sun = -north * 0.5 + up * 0.8660254
    f.globals["cam_forward"] = [0.0, 0.0, -1.0, math.tan(math.radians(4.0))]
    f.globals["cam_up"] = [0.0, 1.0, 0.0, 0.0]
    f.globals["ground_up"] = [*up, 6371000.0]
    f.globals["atmo_center"] = [*center, 6371.0 * km]
    f.bodies = scenes.bodies(
        scenes.body(center, 6371.0 * km, [0.0] * 3, False),
        scenes.body(sun, 0.00465, [115384.68, 109615.445, 101538.516], True),
        scenes.body(sun * (100_000.0 * km) + np.array([1.0, 0.0, 0.0]) * (step * 0.5 * km), 513.0 * km, [0.0] * 3,
                    False),
    )
    return f


def mean_luminance(image):
    """Mean Rec. 709 luminance (src/motion_noise_tests.rs:88-90)."""
    p = image.reshape(-1, 4).astype(np.float64)
    return float(np.mean(0.2126 * p[:, 0] + 0.7152 * p[:, 1] + 0.0722 * p[:, 2]))


def eclipse_energy(make, make_reference, passes=64):
    """The planetshine roulette keeps the expected energy: 4096 spp (64 passes
    of 64) of the production estimator against the same count of an
    estimator that never thins planetshine, on the eclipsed-planetshine and
    eclipsed-sun views, within 3% in mean luminance
    (src/motion_noise_tests.rs:361-405)."""
    size = (32, 18)
    o = dict(max_bounces=1, samples_per_frame=64, sample_limit=4096)
    out = {}
    for name, frame in (("planetshine", eclipsed_planetshine(size, 0)), ("sun", eclipsed_sun(size, 0))):
        expected = mean_luminance(sample(make_reference(size, opts(**o)), frame, passes))
        actual = mean_luminance(sample(make(size, opts(**o)), frame, passes))
        assert expected > 1e-6, (name, expected)
        assert abs(actual / expected - 1.0) < 0.03, (name, actual, expected)
        out[name] = (actual, expected)
    return out


# -- src/noise_tests.rs:5-73 (gpu_ground_noise_is_continuous_at_integer_planes)
def noise_points():
    """Probe points for the value-noise field: pairs straddling integer
    lattice planes along each axis (offsets of +-1e-3, negative coordinates
    included), the lattice points themselves, and metre-scale ground
    patches of the reference's field q = 0.007 p + (19, 2, 4) around places
    where q crosses integer planes."""
    from stargaze_amd.layouts import PROBE_RAY

    rng = np.random.default_rng(73)
    pts = []
    for axis in range(3):
        base = rng.uniform(-60.0, 60.0, (400, 3))
        base[:, axis] = np.round(base[:, axis])
        lo, hi = base.copy(), base.copy()
        lo[:, axis] -= 1e-3
        hi[:, axis] += 1e-3
        pts += [lo, hi]
    lattice = np.round(rng.uniform(-60.0, 60.0, (400, 3)))
    pts.append(lattice)
    # Ground patches 4 m across (the reference's view from 2 m at 70 deg
    # down) centred where 0.007 x + 19 or 0.007 z + 4 is an integer.
    patches = []
    for cx, cz in ((1000.0 / 7.0, 30.0), (-5.0, 1000.0 / 7.0 * 3.0 - 4.0 / 0.007), (-1000.0 / 7.0, -1000.0 / 7.0)):
        g = np.linspace(-2.0, 2.0, 17)
        x, z = np.meshgrid(cx + g, cz + g)
        p = np.stack([x.ravel(), np.zeros(x.size), z.ravel()], axis=1)
        patches.append(0.007 * p + np.array([19.0, 2.0, 4.0]))
    pts += patches
    q = np.concatenate(pts).astype(np.float32)
    rays = np.zeros(len(q), PROBE_RAY)
    rays["anchor"] = 0xFFFFFFFF
    rays["direction"] = q
    rays["target"] = 6  # fbm octaves
    return rays, [len(p) for p in pts], len(patches)


def noise_continuity(probe):
    """Value noise is continuous and stays in [0, 1]: across every integer
    lattice plane the field changes by at most its Lipschitz bound (the
    smoothstep slope 1.5 per lattice unit per axis, so 3e-3 over the 2e-3
    straddle); at lattice points it equals the corner hash; one-octave fbm is
    the value noise itself; and a 4 m ground patch of the reference's field
    (0.028 lattice units across) spans less than 1.5 * 0.028 * 2 = 0.084,
    the bound its metre-scale view relies on (src/noise_tests.rs:63-72)."""
    rays, sizes, n_patches = noise_points()
    out = probe(rays)
    v = out["distance"].astype(np.float64)
    h = out["normal"][:, 0].astype(np.float64)
    f = out["normal"][:, 1].astype(np.float64)
    assert np.all(np.isfinite(v)) and np.all((v >= 0.0) & (v <= 1.0)), (v.min(), v.max())
    assert np.all((h >= 0.0) & (h < 1.0)) and np.all((f >= 0.0) & (f <= 1.0))
    at = np.cumsum([0] + sizes)
    jumps = []
    for axis in range(3):
        lo = v[at[2 * axis]:at[2 * axis + 1]]
        hi = v[at[2 * axis + 1]:at[2 * axis + 2]]
        jumps.append(np.abs(hi - lo).max())
    assert max(jumps) <= 3.0e-3 * 1.01 + 1e-6, jumps
    lattice = slice(at[6], at[7])
    assert np.array_equal(v[lattice], h[lattice]), "vnoise at lattice points is not the corner hash"
    one = rays.copy()
    one["target"] = 1
    assert np.array_equal(probe(one)["normal"][:, 1], out["distance"]), "fbm(q, 1) != vnoise(q)"
    spans = [float(np.ptp(v[at[7 + k]:at[8 + k]])) for k in range(n_patches)]
    assert max(spans) < 0.084, spans
    return max(jumps), spans


# -- src/ground_tests.rs:905-1049 (gpu_forest_skylight_is_occluded_and_has_no_unlit_glow)
def _sky_probe(t, frame, up, count):
    from stargaze_amd.layouts import PROBE_RAY
    from stargaze_amd.pathtracer import PROBE_SKY

    rays = np.zeros(count, PROBE_RAY)
    rays["anchor"] = 0xFFFFFFFF
    rays["direction"] = up
    rays["target"] = np.arange(count)
    return t.probe(frame, PROBE_SKY, rays)


def forest_skylight(make):
    """The surface sky cache and its occlusion, on the lookout in daylight:
    the 16 x 8 cache is finite, non-negative and lit (> 0.01 summed); a
    continued accumulation leaves it unchanged; 64 cosine sky samples from
    the observer towards the zenith gather light (> 1e-3), and under an
    opaque canopy 3 m above the observer, with the same cache, less than a
    thousandth of it (visibility darkens it, not an ambient gain); airless,
    or with every star dark, the cache is zero (< 1e-12)."""
    from stargaze_amd import scenes

    day, data, cells = scenes.lookout(1, 1)
    up = day.globals["ground_up"][:3].astype(np.float64)
    t = make((1, 1), opts(samples_per_frame=1, max_bounces=1))
    t.upload_terrain(data, cells)
    sample(t, day, 1)
    cache = _sky_probe(t, day, up, 128)["extra"].astype(np.float64)
    assert np.all(np.isfinite(cache)) and np.all(cache >= 0.0)
    assert cache.sum() > 0.01, cache.sum()
    sample(t, day, 1)
    assert np.array_equal(_sky_probe(t, day, up, 128)["extra"].astype(np.float64), cache)
    open_sky = float(_sky_probe(t, day, up, 64)["normal"].astype(np.float64).sum())
    assert open_sky > 0.001, open_sky
    covered = scenes.Frame(day.globals.copy(), day.bodies.copy(), day.ground.copy())
    covered.ground[0] = np.zeros((), covered.ground.dtype)
    covered.ground[0]["center"] = [0.0, 5.0, 0.0, 0.0]
    covered.ground[0]["extent"] = [10_000.0, 0.5, 10_000.0, 0.0]
    shaded = float(_sky_probe(t, covered, up, 64)["normal"].astype(np.float64).sum())
    assert shaded < open_sky * 0.001, (shaded, open_sky)
    for airless in (False, True):
        dark = scenes.Frame(day.globals.copy(), day.bodies.copy(), day.ground.copy())
        if airless:
            dark.globals["atmo_params"][3] = 0.0
        else:
            lit = dark.bodies["emissive"] > 0.5
            dark.bodies["color"][lit] = 0.0
        sample(t, dark, 1)
        texels = _sky_probe(t, dark, up, 128)["extra"].astype(np.float64)
        assert np.all(np.isfinite(texels)) and np.all(np.abs(texels) < 1e-12), (airless, np.abs(texels).max())
    return cache.sum(), open_sky, shaded


ALL.append(forest_skylight)


# -- ringshine (src/shaders/pathtrace.wgsl:1236-1296): no reference test; closed form
def _ring_bsdf_factor(tau, mo, mi):
    """Finite-slab single scattering factor of ring_bsdf (src/shaders/rings.wgsl:151-166),
    float64: reflection on the lit side, transmission through the slab."""
    amo, ami = np.maximum(np.abs(mo), 1e-7), np.maximum(np.abs(mi), 1e-7)
    a, b = tau / amo, tau / ami
    same = mo * mi >= 0.0
    refl = (1.0 - np.exp(-(a + b))) / (amo + ami)
    delta = np.abs(a - b)
    integral = np.where(delta < 1e-9, 1.0, (1.0 - np.exp(-delta)) / np.maximum(delta, 1e-300))
    trans = tau / (amo * ami) * np.exp(-np.minimum(a, b)) * integral
    return np.where(same, refl, trans)


def ringshine_frame(tau=0.6, albedo=0.5):
    """A planet (R = 1e-4 AU) with an unbanded ring (1.2-2.3 R, optical depth
    ``tau``) seen through a one-pixel 2e-7 rad view at a night-side point at
    40 deg ring latitude that sees the ring's sunlit face; the sun 1 AU away
    (radius 4e-3 AU) lights the ring from above. Returns the frame and the
    geometry for the quadrature."""
    R = 1e-4
    t = 1e-3
    lat, lon = math.radians(40.0), math.radians(60.0)
    m = np.array([math.cos(lat) * math.cos(lon), math.sin(lat), math.cos(lat) * math.sin(lon)])
    p = np.array([0.0, 0.0, -t])
    c = p - R * m
    s = np.array([-0.6, 0.35, -0.7])
    s /= np.linalg.norm(s)
    n = np.array([0.0, 1.0, 0.0])
    planet = scenes.body(c, R, [albedo] * 3, False)
    planet["ring_plane"] = [*n, 1.2]
    planet["ring_params"] = [2.3, tau, 0.0, 0.0]
    planet["ring_color"] = [0.8, 0.7, 0.6, 0.0]
    planet["axis_x"] = [1.0, 0.0, 0.0, 0.0]
    planet["axis_y"] = [0.0, 0.0, 1.0, 0.0]
    sun_radiance, sun_radius = 2.0e4, 4e-3
    f = scenes.frame()
    f.bodies = scenes.bodies(planet, scenes.body(s * 1.0, sun_radius, [sun_radiance] * 3, True))
    geometry = dict(R=R, c=c, p=p, m=m, s=s, n=n, tau=tau, albedo=albedo, color=np.array([0.8, 0.7, 0.6]),
                    sun=sun_radiance, omega=2.0 * math.pi * (1.0 - math.sqrt(1.0 - sun_radius ** 2)))
    return f, geometry


def ringshine_expected(g, radial=1500, around=3000):
    """E[pixel] = albedo / pi * integral over the ring annulus of the ring's
    single-scattered sunlight towards p, cos(surface) cos(ring) / d^2, with
    the planet's shadow on the ring and its own horizon at p: midpoint
    quadrature in radius and azimuth, float64."""
    R, c, p, m, s, n = g["R"], g["c"], g["p"], g["m"], g["s"], g["n"]
    r = R * (1.2 + 1.1 * (np.arange(radial) + 0.5) / radial)
    phi = 2.0 * math.pi * (np.arange(around) + 0.5) / around
    rr, pp = np.meshgrid(r, phi, indexing="ij")
    e1, e2 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    x = c + rr[..., None] * (np.cos(pp)[..., None] * e1 + np.sin(pp)[..., None] * e2)
    v = x - p
    d = np.linalg.norm(v, axis=-1)
    u = v / d[..., None]
    cos_s = u @ m
    cos_r = np.abs(u @ n)
    mo = -(u @ n)
    mi = float(s @ n)
    # Sun hidden by the planet from x: does x + t s (t > 0) meet the sphere?
    oc = x - c
    b = oc @ s
    disc = b * b - (np.einsum("...i,...i", oc, oc) - R * R)
    shadow = (disc > 0.0) & (-b - np.sqrt(np.maximum(disc, 0.0)) > 0.0)
    term = g["color"][None, None, :] * (_ring_bsdf_factor(g["tau"], mo, mi) / (4.0 * math.pi))[..., None] * \
        (g["sun"] * g["omega"] * abs(mi))
    weight = np.where((cos_s > 0.0) & (cos_r >= 1e-6) & ~shadow, cos_s * cos_r / (d * d), 0.0)
    area = rr * (1.1 * R / radial) * (2.0 * math.pi / around)
    return g["albedo"] / math.pi * np.einsum("ijk,ij->k", term, weight * area)


def ringshine_closed_form(make):
    """Ringshine, the one light source of a night-side surface point: a
    million samples at the last vertex (ringshine at full weight) against
    the quadrature of its integral, per channel, within 4 standard errors of
    the estimate and 1%. The ring's radiance, its area pdf and the cosines all
    enter, so a dropped term or a wrong sampling density fails it."""
    f, g = ringshine_frame()
    expected = ringshine_expected(g)
    # 32 x 32 pixels of the 2e-7 rad view all see p (to 1e-4 of R): a
    # million samples in all.
    t = make((32, 32), opts(samples_per_frame=64, max_bounces=1))
    t.enable_moments(True)
    pixels = sample(t, f, 16)[:, :3].astype(np.float64)
    m2 = t.read_moments().reshape(-1, 4).astype(np.float64)
    mean = pixels.mean(axis=0)
    n = m2[:, 3].sum()
    variance = np.mean(m2[:, :3] - pixels * pixels, axis=0) + pixels.var(axis=0)
    se = np.sqrt(np.maximum(variance, 0.0) / n)
    assert np.all(expected > 0.0), expected
    assert np.all(np.abs(mean - expected) <= 4.0 * se + 0.01 * expected), (mean, expected, se)
    return mean, expected, se


ALL.append(ringshine_closed_form)
