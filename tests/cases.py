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
]
