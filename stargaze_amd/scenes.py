"""Scene builders: the reference's test frames and the benchmark workloads.

Frames follow the reference's conventions (src/main.rs:822-1018): telescope
space looks along -Z, lengths in AU, body centres split into an f32 high part
and an f32 residual, star radiance ``2.5 * luminosity / radius_au**2``
(src/main.rs:71,846). Workloads are synthetic, seeded and self-contained.
"""
from __future__ import annotations

import math

import numpy as np

from .layouts import AU_KM, GLOBALS, PRIMITIVE, SPHERE
from .pathtracer import Frame

METRES_PER_AU = AU_KM * 1000.0
SUN_LIGHT = 2.5  # src/main.rs:71


def metres(x: float) -> float:
    """Metres to AU."""
    return x / METRES_PER_AU


# -- the reference's unit-test helpers (src/pathtracer_tests.rs:32-61,
#    src/ring_tests.rs:4-10) -------------------------------------------------
def frame() -> Frame:
    g = np.zeros((), GLOBALS)
    g["cam_right"] = [1, 0, 0, 0]
    g["cam_up"] = [0, 1, 0, 0]
    g["cam_forward"] = [0, 0, -1, 1e-7]
    g["viewport"] = [1, 1, 1, 0]
    return Frame(globals=g)


def body(center, radius, color, emissive: bool) -> np.ndarray:
    b = np.zeros((), SPHERE)
    b["center"] = center
    b["radius"] = radius
    b["color"] = color
    b["emissive"] = 1.0 if emissive else 0.0
    return b


def ringed(center, radius, pole, tau) -> np.ndarray:
    planet = body(center, radius, [0.0] * 3, False)
    planet["ring_plane"] = [pole[0], pole[1], pole[2], 1.25]
    planet["ring_params"] = [2.4, tau, 0.0, 0.0]
    planet["ring_color"] = [0.8, 0.8, 0.8, 0.0]
    return planet


def bodies(*records) -> np.ndarray:
    return np.array([np.asarray(r, SPHERE) for r in records], SPHERE).reshape(-1)


# -- helpers for physical scenes ---------------------------------------------
def split(center_au: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """High/low f32 split of an f64 telescope-space centre (src/main.rs:838-844)."""
    c = np.asarray(center_au, np.float64)
    hi = c.astype(np.float32)
    lo = (c - hi.astype(np.float64)).astype(np.float32)
    return hi, lo


def place(b: np.ndarray, center_au) -> np.ndarray:
    hi, lo = split(center_au)
    b["center"] = hi
    b["center_low"][:3] = lo
    return b


def star(center_au, radius_au: float, luminosity: float = 1.0) -> np.ndarray:
    s = body([0, 0, 0], radius_au, [SUN_LIGHT * luminosity / radius_au**2] * 3, True)
    return place(s, center_au)


def direction(azimuth_deg: float, elevation_deg: float) -> np.ndarray:
    """Unit vector in a frame with +Y up and -Z forward (azimuth to +X)."""
    a, e = math.radians(azimuth_deg), math.radians(elevation_deg)
    return np.array([math.cos(e) * math.sin(a), math.sin(e), -math.cos(e) * math.cos(a)])


def camera(g: np.ndarray, fov_deg: float, width: int, height: int) -> None:
    g["cam_right"] = [1, 0, 0, 0]
    g["cam_up"] = [0, 1, 0, 0]
    g["cam_forward"] = [0, 0, -1, math.tan(math.radians(fov_deg) * 0.5)]
    g["viewport"] = [width, height, 1.0, 0.0]


class Pcg:
    """Seeded PCG32 for workload synthesis (same step as the path RNG)."""

    def __init__(self, seed: int):
        self.state = seed & 0xFFFFFFFF

    def next_u32(self) -> int:
        self.state = (self.state * 747796405 + 2891336453) & 0xFFFFFFFF
        s = self.state
        word = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
        return ((word >> 22) ^ word) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u32() >> 8) / 16777216.0


# -- workloads ---------------------------------------------------------------
def airless_triple(width: int = 256, height: int = 256) -> Frame:
    """Star + planet + moon, camera 2 m above the planet, 60 deg FOV.

    Star R = 7e8 m at 1.5e11 m; planet R = 6.4e6 m, albedo 0.3; moon
    R = 1.7e6 m, albedo 0.12, 3.8e8 m from the planet. Constant albedos
    (procedural strength 0), no rings, air, ground or catalogue.
    """
    f = frame()
    camera(f.globals, 60.0, width, height)
    r_planet = metres(6.4e6)
    up = np.array([0.0, 1.0, 0.0])
    planet_c = -up * (r_planet + metres(2.0))
    sun_dir = direction(-70.0, 25.0)
    moon_dir = direction(12.0, 9.0)
    moon_c = planet_c + moon_dir * metres(3.8e8)
    planet = place(body([0, 0, 0], r_planet, [0.3] * 3, False), planet_c)
    moon = place(body([0, 0, 0], metres(1.7e6), [0.12] * 3, False), moon_c)
    sun = star(sun_dir * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, planet, moon)
    return f


def ringed_giant(width: int = 1920, height: int = 1080, seed: int = 0x5EED1234) -> Frame:
    """Banded gas giant with a C/B/A ring system, four moons, one star.

    Giant R = 6.0e7 m, gas-giant bands between albedo 0.2 and 0.6; moons with
    radii log-uniform in [2e5, 2.6e6] m and albedo U[0.05, 0.8] (PCG seed
    0x5EED1234); equatorial ring 6.7e7-1.4e8 m with the banded optical-depth
    profile (C ~0.1, B ~1.5, A ~0.5, Cassini/Encke/Keeler gaps), single
    scattering albedo 0.6.
    """
    rng = Pcg(seed)
    f = frame()
    camera(f.globals, 50.0, width, height)
    r_giant = metres(6.0e7)
    giant_c = np.array([0.0, 0.0, -metres(3.6e8)])
    giant = body([0, 0, 0], r_giant, [0.6, 0.55, 0.45], False)
    giant["axis_x"] = [1, 0, 0, 1.0]  # style 1: gas giant
    pole = np.array([0.0, math.cos(math.radians(22.0)), math.sin(math.radians(22.0))])
    x_axis = np.array([1.0, 0.0, 0.0])
    y_axis = np.cross(pole, x_axis)
    giant["axis_x"][:3] = x_axis
    giant["axis_y"] = [*y_axis, 7.0]
    giant["surface_color"] = [0.2, 0.17, 0.13, 0.8]
    giant["surface_params"] = [7.0, 2.0, 0.0, 0.0]
    giant["ring_plane"] = [*pole, 6.7e7 / 6.0e7]
    giant["ring_params"] = [1.4e8 / 6.0e7, 1.0, 1.0, 0.0]
    giant["ring_color"] = [0.6, 0.6, 0.6, 0.0]
    place(giant, giant_c)
    moons = []
    for k in range(4):
        radius = metres(math.exp(rng.uniform(math.log(2e5), math.log(2.6e6))))
        albedo = rng.uniform(0.05, 0.8)
        angle = rng.uniform(0.0, 2 * math.pi)
        dist = metres(rng.uniform(1.6e8, 2.6e8))
        lift = rng.uniform(-0.08, 0.08)
        offset = np.array([math.cos(angle) * dist, lift * dist, math.sin(angle) * dist * 0.6])
        m = place(body([0, 0, 0], radius, [albedo] * 3, False), giant_c + offset)
        moons.append(m)
    sun = star(direction(-50.0, 30.0) * 9.5, metres(7.0e8))
    f.bodies = bodies(sun, giant, *moons)
    return f


# Style codes of body.axis_x.w (src/sim.rs:109-118).
STYLE_PLAIN, STYLE_GAS_GIANT, STYLE_CRATERED, STYLE_ROCKY, STYLE_ICY, STYLE_VOLCANIC, STYLE_TERRAN = range(7)


def style_surface(b: np.ndarray, style: int, seed: float = 3.0, craters: float | None = None,
                  ice_caps: float | None = None, secondary=None, ocean: float = 0.65,
                  clouds: float = 0.4) -> np.ndarray:
    """Give body ``b`` a solid-body surface (styles 2-6) with the reference's
    defaults for the style (src/config.rs:216-251): secondary albedo, crater
    density and ice caps per style; ocean 0.65 and clouds 0.4. The record
    layout is the one the shader reads (src/main.rs:872-895): style in
    axis_x.w, seed in axis_y.w, secondary albedo and crater density in
    surface_color, (ice caps, ocean, clouds) in surface_params.xyz."""
    albedo = [float(c) for c in b["color"]]
    peak = max(max(albedo), 1.0e-6)

    def deeper(scale, power):
        return [scale * c * (c / peak) ** power for c in albedo]

    a = albedo
    defaults = {
        STYLE_CRATERED: (deeper(0.55, 0.6), 0.8, 0.0),
        STYLE_ROCKY: (deeper(0.6, 1.0), 0.5, 0.06),
        STYLE_ICY: ([a[0] * 0.75, a[1] * 0.5, a[2] * 0.35], 0.15, 0.0),
        STYLE_VOLCANIC: ([a[0] * 0.85, a[1] * 0.45, a[2] * 0.2], 0.5, 0.0),
        STYLE_TERRAN: ([0.03, 0.09, 0.2], 0.0, 0.1),
    }
    second, crater_default, ice_default = defaults[style]
    if secondary is not None:
        second = list(secondary)
    b["axis_x"][3] = float(style)
    b["axis_y"][3] = float(seed)
    b["surface_color"] = [*second, crater_default if craters is None else craters]
    b["surface_params"] = [ice_default if ice_caps is None else ice_caps, ocean, clouds, 0.0]
    return b


def material_frame(style: int, albedo=(0.62, 0.6, 0.57), seed: float = 3.0, **surface) -> Frame:
    """A frame whose body 1 carries a procedural surface, for the material
    probe (src/surface_tests.rs:61-86, 87-169): star, then the styled body
    with its spin frame along the world axes."""
    f = frame()
    camera(f.globals, 45.0, 64, 64)
    b = body([0, 0, 0], metres(1.7e6), list(albedo), False)
    b["axis_x"][:3] = [1.0, 0.0, 0.0]
    b["axis_y"][:3] = [0.0, 1.0, 0.0]
    if style == STYLE_PLAIN:
        b["axis_y"][3] = float(seed)
        b["center_low"][3] = 0.5
    else:
        style_surface(b, style, seed, **surface)
    place(b, np.array([0.0, 0.0, -metres(1.0e7)]))
    sun = star(direction(-40.0, 20.0) * 9.5, metres(7.0e8))
    f.bodies = bodies(sun, b)
    return f


STYLE_ALBEDOS = {STYLE_CRATERED: (0.62, 0.6, 0.57), STYLE_ROCKY: (0.7, 0.45, 0.3), STYLE_ICY: (0.85, 0.8, 0.72),
                 STYLE_VOLCANIC: (0.85, 0.75, 0.3), STYLE_TERRAN: (0.55, 0.45, 0.3)}


def styled_moons(width: int = 64, height: int = 48, seed: float = 3.0, styles=(STYLE_CRATERED, STYLE_ROCKY)) -> Frame:
    """Two moons with solid-body surfaces side by side (by default a cratered
    one and a rocky one with polar caps), lit by one star behind the camera,
    filling most of the view."""
    f = frame()
    camera(f.globals, 50.0, width, height)
    moons = []
    for style, x in zip(styles, (-1.9e6, 1.9e6)):
        albedo = STYLE_ALBEDOS[style]
        b = body([0, 0, 0], metres(1.7e6), list(albedo), False)
        pole = np.array([0.0, math.cos(0.4), math.sin(0.4)])
        b["axis_x"][:3] = [1.0, 0.0, 0.0]
        b["axis_y"][:3] = np.cross(pole, [1.0, 0.0, 0.0])
        style_surface(b, style, seed, ice_caps=0.25 if style == STYLE_ROCKY else None)
        moons.append(place(b, np.array([metres(x), 0.0, -metres(8.0e6)])))
    sun = star(direction(-130.0, 25.0) * 9.5, metres(7.0e8))
    f.bodies = bodies(sun, *moons)
    return f


def material_rays(width: int = 256, height: int = 128) -> np.ndarray:
    """Unit points of a latitude-longitude grid (src/surface_tests.rs:108-113),
    as probe rays aimed at body 1."""
    from .layouts import PROBE_RAY

    y, x = np.mgrid[0:height, 0:width]
    u = (x + 0.5) / width
    v = (y + 0.5) / height
    lat = (v - 0.5) * math.pi
    lon = u * 2.0 * math.pi
    q = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], -1).reshape(-1, 3)
    rays = np.zeros(q.shape[0], PROBE_RAY)
    rays["direction"] = q.astype(np.float32)
    rays["anchor"] = 0xFFFFFFFF
    rays["target"] = 1
    return rays


WORKLOADS = {
    "airless_triple_256": (airless_triple, dict(width=256, height=256), 16, 4),
    "ringed_giant_1080p": (ringed_giant, dict(width=1920, height=1080), 64, 6),
}


# -- the Halo observatory ----------------------------------------------------
def _ground_frame(pitch_deg: float):
    """(east, up, north) in telescope space for a camera looking north,
    pitched up by ``pitch_deg`` (src/camera.rs:74-91 conventions)."""
    t = math.radians(pitch_deg)
    east = np.array([1.0, 0.0, 0.0])
    up = np.array([0.0, math.cos(t), -math.sin(t)])
    north = np.array([0.0, -math.sin(t), -math.cos(t)])
    return east, up, north


def sky_direction(east, up, north, azimuth_deg, elevation_deg):
    a, e = math.radians(azimuth_deg), math.radians(elevation_deg)
    return math.cos(e) * (math.cos(a) * north + math.sin(a) * east) + math.sin(e) * up


def _rotation_z_extent(extent, angle):
    """``Primitive::world_extent`` for a rotation about the north (z) axis."""
    c, s = abs(math.cos(angle)), abs(math.sin(angle))
    return np.array([c * extent[0] + s * extent[1], s * extent[0] + c * extent[1], extent[2]])


def _group(center_lo, center_hi, count):
    g = np.zeros((), PRIMITIVE)
    g["center"] = [*((center_lo + center_hi) * 0.5), -float(count)]
    g["extent"] = [*((center_hi - center_lo) * 0.5 + 0.02), 0.0]
    return g


def cluster_props(props: np.ndarray, levels=((8, 8), (4, 4), (4, 4))) -> np.ndarray:
    """Nested groups in the reference's flat skip-count format
    (src/ground.rs:198-225, src/shaders/ground.wgsl:167-178): a group marker
    (center.w = -descendants) bounds every following entry it owns; a ray
    that misses it skips them all."""
    centre = props["center"][:, :3].astype(np.float64)
    e = props["extent"].astype(np.float64)
    c, s_ = np.abs(np.cos(e[:, 3])), np.abs(np.sin(e[:, 3]))
    ext = np.stack([c * e[:, 0] + s_ * e[:, 1], s_ * e[:, 0] + c * e[:, 1], e[:, 2]], -1)
    lo_all, hi_all = centre - ext, centre + ext

    def build(idx, depth, x0, x1, z0, z1):
        if depth == len(levels):
            return list(props[idx])
        nx, nz = levels[depth]
        out = []
        xs = np.linspace(x0, x1, nx + 1)
        zs = np.linspace(z0, z1, nz + 1)
        cx = np.clip(np.searchsorted(xs, centre[idx, 0], side="right") - 1, 0, nx - 1)
        cz = np.clip(np.searchsorted(zs, centre[idx, 2], side="right") - 1, 0, nz - 1)
        for j in range(nz):
            for i in range(nx):
                sel = idx[(cx == i) & (cz == j)]
                if sel.size == 0:
                    continue
                children = build(sel, depth + 1, xs[i], xs[i + 1], zs[j], zs[j + 1])
                single_group = children[0]["center"][3] < 0 and -children[0]["center"][3] == len(children) - 1
                if len(children) == 1 or single_group:
                    out.extend(children)
                    continue
                lo = lo_all[sel].min(axis=0)
                hi = hi_all[sel].max(axis=0)
                out.append(_group(lo, hi, len(children)))
                out.extend(children)
        return out

    x0, z0 = centre[:, 0].min() - 1, centre[:, 2].min() - 1
    x1, z1 = centre[:, 0].max() + 1, centre[:, 2].max() + 1
    flat = build(np.arange(len(props)), 0, x0, x1, z0, z1)
    return np.array(flat, PRIMITIVE)


def scatter_props(heights, cells, radius_m, rocks, conifers, seed=0x1234ABCD, reach=2500.0, clear=12.0):
    """Boxes (rocks, stone material) and ellipsoids (conifers, foliage) on
    the terrain surface, uniform in area between ``clear`` and ``reach`` m."""
    from . import terrain as tr

    rng = np.random.default_rng(seed)
    s = tr.scale(radius_m)
    n = rocks + conifers
    r = np.sqrt(rng.uniform(clear**2, reach**2, n))
    a = rng.uniform(0, 2 * math.pi, n)
    xs, zs = r * np.cos(a), r * np.sin(a)
    d2 = xs * xs + zs * zs
    curvature = -d2 / (radius_m + np.sqrt(radius_m * radius_m - d2))
    ground = tr.heights_at(heights, cells, xs / s, zs / s) * s + curvature
    props = np.zeros(n, PRIMITIVE)
    e = rng.uniform(0.25, 1.4, (rocks, 3)) * np.array([1.0, 0.6, 1.0])
    angle = rng.uniform(-0.35, 0.35, rocks)
    grey = rng.uniform(0.18, 0.34, rocks)
    props["center"][:rocks] = np.stack([xs[:rocks], ground[:rocks] + e[:, 1] * 0.6, zs[:rocks], np.zeros(rocks)], -1)
    props["extent"][:rocks] = np.concatenate([e, angle[:, None]], -1)
    props["albedo"][:rocks] = np.stack([grey, grey * 0.97, grey * 0.92, np.full(rocks, 2.0)], -1)
    rad = rng.uniform(1.2, 2.8, conifers)
    half = rng.uniform(4.0, 11.0, conifers)
    green = rng.uniform(0.09, 0.13, conifers)
    c = slice(rocks, n)
    props["center"][c] = np.stack([xs[c], ground[c] + half * 0.92, zs[c], np.ones(conifers)], -1)
    props["extent"][c] = np.stack([rad, half, rad, np.zeros(conifers)], -1)
    props["albedo"][c] = np.stack([np.full(conifers, 0.045), green, np.full(conifers, 0.035), np.full(conifers, 3.0)], -1)
    return props


def halo_observatory(width=3840, height=2160, cells=4096, rocks=20000, conifers=50000, seed=0x743337A9,
                     reach=2500.0, levels=((8, 8), (4, 4), (4, 4))):
    """Halo observatory: the camera 2 m above a 1.2e6 m host at 35 deg N,
    ridged-fBm terrain, rocks and conifers, Earth-like air with Rayleigh
    beta (5.8, 13.5, 33.1)e-6 /m (8 km scale height) and Mie 2e-5 /m
    (g = 0.76), two inflated stars low in the sky and the ringed giant
    overhead. The frame carries one scale height, as the reference's
    Globals do; the 4K workload's 16 x 8 march and 1.2 km Mie scale height
    are the tracer settings of ``HALO_AIR`` (``PathTracer.set_air_model``).

    Returns ``(frame, terrain_data, cells)``; upload the terrain with
    ``PathTracer.upload_terrain(terrain_data, cells)`` before encoding.
    """
    from . import terrain as tr

    radius_m = 1.2e6
    f = frame()
    camera(f.globals, 70.0, width, height)
    east, up, north = _ground_frame(15.0)
    r_host = metres(radius_m)
    host_c = -up * (r_host + metres(2.0))
    host = place(body([0, 0, 0], r_host, [0.25, 0.22, 0.18], False), host_c)
    sun_a = star(sky_direction(east, up, north, -40.0, 27.0) * 2.0, metres(2.2e9), 3.0)
    sun_b = star(sky_direction(east, up, north, -49.0, 31.0) * 2.0, metres(1.3e9), 1.2)
    giant_frame = ringed_giant(width, height)
    giant = giant_frame.bodies[1].copy()
    giant_c = sky_direction(east, up, north, 12.0, 40.0) * metres(1.6e9)
    pole = up * 0.8 + north * 0.6
    pole /= np.linalg.norm(pole)
    giant["ring_plane"][:3] = pole
    x_axis = np.cross(pole, east)
    x_axis /= np.linalg.norm(x_axis)
    giant["axis_x"][:3] = x_axis
    giant["axis_y"][:3] = np.cross(pole, x_axis)
    place(giant, giant_c)
    f.bodies = bodies(sun_a, sun_b, host, giant)
    host_index = 2
    au = METRES_PER_AU
    g = f.globals
    g["atmo_center"] = [*split(host_c)[0], r_host]
    g["atmo_rayleigh"] = [5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au, 2.0e-5 * au]
    g["atmo_params"] = [0.76, 8000.0 / au, 80000.0 / au, host_index + 1]
    g["ground_east"] = [*east, 2.0]
    g["ground_up"] = [*up, radius_m]
    g["ground_north"] = [*north, host_index + 1]
    heights = tr.ridged_fbm(cells, seed=seed, amplitude=800.0)
    data = tr.terrain_data(heights, cells)
    props = scatter_props(heights, cells, radius_m, rocks, conifers, reach=reach)
    ground = cluster_props(props, levels) if len(props) > 64 else props
    f.ground = ground
    g["ground_counts"] = [len(ground), 1, 0, 0]
    return f, data, cells


# Air model of the Halo 4K workload: 16 view cells per camera ray, 8 density
# samples per light path, Mie scale height 1.2 km (Rayleigh keeps 8 km).
HALO_AIR = dict(view_steps=16, light_steps=8, mie_scale_height=1200.0 / METRES_PER_AU)


def lookout(width=96, height=54, theme=0, radius_m=6.371e6, cells=256, seed=0x743337A9, pitch_deg=4.0):
    """The reference's own ground scenery around a lookout 2 m above an
    Earth-sized host with Earth-like air and one sun at 30 deg: the forest
    clearing of ``ground.landscape_with_density`` (stone outcrops, conifers,
    undergrowth, logs and the cottage: every primitive kind but the tumbled
    rock) for theme 0, or the tumbled rocks of ``ground.theme_rocks`` for
    the desert (1), crater (2) and basalt (4) grounds, on a ridged-fBm
    terrain whose heights the props stand on. The camera looks north along
    the observing gap, pitched down ``pitch_deg``. Returns
    ``(frame, terrain_data, cells)``."""
    from . import ground as gr
    from . import terrain as tr

    f = frame()
    camera(f.globals, 70.0, width, height)
    east, up, north = _ground_frame(-pitch_deg)
    r_host = metres(radius_m)
    host_c = -up * (r_host + metres(2.0))
    host = place(body([0, 0, 0], r_host, [0.25, 0.22, 0.18], False), host_c)
    sun = star(sky_direction(east, up, north, 35.0, 30.0), metres(7.0e8))
    f.bodies = bodies(sun, host)
    au = METRES_PER_AU
    g = f.globals
    g["atmo_center"] = [*split(host_c)[0], r_host]
    g["atmo_rayleigh"] = [5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au, 2.0e-5 * au]
    g["atmo_params"] = [0.76, 8000.0 / au, 80000.0 / au, 2.0]
    g["ground_east"] = [*east, 2.0]
    g["ground_up"] = [*up, radius_m]
    g["ground_north"] = [*north, 2.0]
    heights = tr.ridged_fbm(cells, seed=seed, amplitude=60.0)
    s = tr.scale(radius_m)
    surface = lambda x, z: float(tr.heights_at(heights, cells, np.array([x / s]), np.array([z / s]))[0]) * s  # noqa: E731
    if theme == 0:
        f.ground = gr.landscape_with_density(radius_m, 100, surface)
    else:
        f.ground = gr.theme_rocks(radius_m, theme, surface)
    g["ground_counts"] = [len(f.ground), 1, theme, 0]
    return f, tr.terrain_data(heights, cells), cells


def halo_parity(width=96, height=54):
    """Halo scene inside the reference's capacities: 512-cell terrain and at
    most 256 primitives (src/terrain.rs:20, src/ground.rs:8)."""
    return halo_observatory(width, height, cells=512, rocks=50, conifers=130, reach=400.0, levels=((6, 6),))


# -- batched views -------------------------------------------------------------
def orbit_air(width: int = 320, height: int = 240, seed: int = 0x0A11CE) -> Frame:
    """A 6.4e6 m planet with Earth-like air seen from 4e6 m above its
    surface, the limb across the view, one star; no ground scene."""
    rng = Pcg(seed)
    f = frame()
    camera(f.globals, 55.0, width, height)
    r = metres(6.4e6)
    # Planet centre 35-50 deg below the view axis: from 4e6 m up its limb
    # (38 deg from the centre) crosses the view.
    below = math.radians(rng.uniform(35.0, 50.0))
    centre = np.array([0.0, -math.sin(below), -math.cos(below)]) * (r + metres(4.0e6))
    planet = place(body([0, 0, 0], r, [0.28, 0.3, 0.33], False), centre)
    sun = star(direction(rng.uniform(-90.0, 90.0), rng.uniform(5.0, 40.0)) * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, planet)
    au = METRES_PER_AU
    g = f.globals
    g["atmo_center"] = [*split(centre)[0], r]
    g["atmo_rayleigh"] = [5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au, 2.0e-5 * au]
    g["atmo_params"] = [0.76, 8000.0 / au, 80000.0 / au, 2.0]
    return f


def batch_views(count: int = 96, seed: int = 0xBA7C4, min_pixels: int = 128 * 128, max_pixels: int = 1920 * 1080,
                spp=(1, 2, 4, 8, 16), bounces: int = 4, halo_every: int = 6, terrain_cells: int = 512):
    """``count`` views of uneven resolution and sample count: pixel counts
    log-uniform in [min_pixels, max_pixels] at aspect 1, 4:3 or 16:9,
    samples per pass drawn from ``spp``, scenes cycling through the airless
    triple, the ringed giant (own moon seed), a planet with air seen from
    orbit and, every ``halo_every``-th view, the Halo observatory at the
    reference's terrain and primitive capacities.

    Returns a list of dicts with ``frame``, ``terrain`` (``(data, cells)`` or
    None), ``width``, ``height``, ``spp`` and ``bounces``.
    """
    rng = np.random.default_rng(seed)
    aspects = (1.0, 4.0 / 3.0, 16.0 / 9.0)
    halo_cache = {}
    views = []
    for i in range(count):
        pixels = math.exp(rng.uniform(math.log(min_pixels), math.log(max_pixels)))
        aspect = aspects[int(rng.integers(len(aspects)))]
        height = max(8, int(round(math.sqrt(pixels / aspect))))
        width = max(8, int(round(height * aspect)))
        samples = int(spp[int(rng.integers(len(spp)))])
        terrain = None
        if halo_every and i % halo_every == halo_every - 1:
            if "halo" not in halo_cache:
                f, data, cells = halo_observatory(64, 36, cells=terrain_cells, rocks=50, conifers=130, reach=400.0,
                                                  levels=((6, 6),))
                halo_cache["halo"] = (f, data, cells)
            f0, data, cells = halo_cache["halo"]
            f = Frame(f0.globals.copy(), f0.bodies.copy(), f0.ground.copy())
            camera(f.globals, 70.0, width, height)
            terrain = (data, cells)
            kind = "halo"
        else:
            kind = ("triple", "giant", "orbit_air")[i % 3]
            if kind == "triple":
                f = airless_triple(width, height)
            elif kind == "giant":
                f = ringed_giant(width, height, seed=int(rng.integers(1, 2**31)))
            else:
                f = orbit_air(width, height, seed=int(rng.integers(1, 2**31)))
        views.append(dict(kind=kind, frame=f, terrain=terrain, width=width, height=height, spp=samples,
                          bounces=bounces))
    return views


# -- differentiable-rendering scenes ---------------------------------------------
def moons_over_plain(width: int = 48, height: int = 32) -> Frame:
    """Star, a planet under the camera (2 m up) and three close moons high
    in a 60 deg view, the third with procedural mottle; albedos in (0, 1)."""
    f = frame()
    camera(f.globals, 60.0, width, height)
    r_planet = metres(6.4e6)
    up = np.array([0.0, 1.0, 0.0])
    planet_c = -up * (r_planet + metres(2.0))
    planet = place(body([0, 0, 0], r_planet, [0.31, 0.27, 0.22], False), planet_c)
    moons = []
    specs = [((-14.0, 14.0), 2.4e6, 2.0e7, [0.55, 0.5, 0.45]), ((6.0, 20.0), 1.6e6, 1.4e7, [0.2, 0.35, 0.6]),
             ((18.0, 9.0), 2.0e6, 1.8e7, [0.7, 0.62, 0.4])]
    for k, ((az, el), radius, dist, color) in enumerate(specs):
        m = place(body([0, 0, 0], metres(radius), color, False), direction(az, el) * metres(dist))
        if k == 2:
            m["center_low"][3] = 0.8  # procedural mottle strength
            m["axis_x"] = [1, 0, 0, 0.0]
            m["axis_y"] = [0, 1, 0, 3.0]
        moons.append(m)
    sun = star(direction(-60.0, 35.0) * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, planet, *moons)
    return f


def ring_shadow(width: int = 48, height: int = 32, tau: float = 0.6, outer: float = 2.3) -> Frame:
    """A ringed giant seen from 1.5 radii above 60 deg latitude, the view
    filled by its northern hemisphere; the star lies 30 deg below the ring
    plane, on the camera's side, so shadow rays from the lit surface meet
    the ring (1.1 to ``outer`` radii) near two radii. Camera rays never
    reach the ring plane."""
    f = frame()
    camera(f.globals, 40.0, width, height)
    r = metres(6.0e7)
    lat = math.radians(60.0)
    # Camera frame: forward -z; the planet centre straight ahead.
    centre = np.array([0.0, 0.0, -1.5 * r])
    pole = np.array([0.0, math.cos(lat), math.sin(lat)])  # n . (camera - centre) / |..| = sin(lat)
    giant = body([0, 0, 0], r, [0.5, 0.45, 0.4], False)
    giant["ring_plane"] = [*pole, 1.1]
    giant["ring_params"] = [outer, tau, 0.0, 0.0]
    giant["ring_color"] = [0.7, 0.65, 0.6, 0.0]
    x_axis = np.array([1.0, 0.0, 0.0])
    giant["axis_x"] = [*x_axis, 0.0]
    giant["axis_y"] = [*np.cross(pole, x_axis), 0.0]
    place(giant, centre)
    # In-plane direction towards the camera's sub-point, then the star:
    # 30 deg below the ring plane, on the camera's side.
    towards = np.array([0.0, 0.0, 1.0]) - pole * math.sin(lat)
    towards /= np.linalg.norm(towards)
    light = towards * math.cos(math.radians(30.0)) - pole * math.sin(math.radians(30.0))
    sun = star(light * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, giant)
    return f


def ring_window(width: int = 32, height: int = 32, tau: float = 0.5) -> Frame:
    """Looking straight through a black (colour 0) ring at a star that fills
    the 1 deg view: every camera ray crosses the ring annulus near normal
    incidence, so a pixel's expected radiance is L exp(-tau / mu)."""
    f = frame()
    camera(f.globals, 1.0, width, height)
    r = metres(6.0e7)
    pole = np.array([0.0, 0.0, 1.0])
    giant = body([0, 0, 0], r, [0.5, 0.5, 0.5], False)
    giant["ring_plane"] = [*pole, 1.1]
    giant["ring_params"] = [2.3, tau, 0.0, 0.0]
    giant["ring_color"] = [0.0, 0.0, 0.0, 0.0]
    # Ring centre off to the side so the line of sight crosses at 1.7 radii.
    place(giant, np.array([1.7 * r, 0.0, -5.0 * r]))
    sun = star(np.array([0.0, 0.0, -1.0]) * metres(1.5e11), metres(7.0e9))
    f.bodies = bodies(sun, giant)
    return f


def ringed_disc(width: int = 64, height: int = 48, tau: float = 0.6) -> Frame:
    """A ringed giant filling half of a 40 deg view from four radii, ring
    plane tilted 20 deg from edge-on, lit from above and behind the camera:
    the lit ring crosses the disc and its shadow falls on the disc below it."""
    f = frame()
    camera(f.globals, 40.0, width, height)
    r = metres(6.0e7)
    centre = np.array([0.0, 0.0, -4.0 * r])
    tilt = math.radians(20.0)
    pole = np.array([0.0, math.cos(tilt), math.sin(tilt)])
    giant = body([0, 0, 0], r, [0.5, 0.45, 0.4], False)
    giant["ring_plane"] = [*pole, 1.2]
    giant["ring_params"] = [2.2, tau, 0.0, 0.0]
    giant["ring_color"] = [0.75, 0.7, 0.62, 0.0]
    x_axis = np.array([1.0, 0.0, 0.0])
    giant["axis_x"] = [*x_axis, 0.0]
    giant["axis_y"] = [*np.cross(pole, x_axis), 0.0]
    place(giant, centre)
    light = np.array([0.25, 0.55, 1.0])
    light /= np.linalg.norm(light)
    sun = star(light * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, giant)
    return f


def diff_system(width: int = 512, height: int = 512, seed: int = 2) -> Frame:
    """The differentiable-rendering workload: from 3.0e6 m above a 6.4e6 m
    planet with Earth-like air, its limb across the lower view, a ringed
    giant (constant-albedo surface, ring 1.2-2.3 radii) and six moons in the
    sky, one star. Eight lit bodies with style-0 albedos (mottled moons),
    their albedos drawn with a PCG seeded by ``seed``."""
    rng = Pcg(seed)
    f = frame()
    camera(f.globals, 60.0, width, height)
    r = metres(6.4e6)
    below = math.radians(48.0)
    host_c = np.array([0.0, -math.sin(below), -math.cos(below)]) * (r + metres(3.0e6))
    host = place(body([0, 0, 0], r, [0.3, 0.32, 0.35], False), host_c)
    r_giant = metres(6.0e7)
    giant = body([0, 0, 0], r_giant, [0.6, 0.52, 0.4], False)
    tilt = math.radians(18.0)
    pole = np.array([0.0, math.cos(tilt), math.sin(tilt)])
    giant["ring_plane"] = [*pole, 1.2]
    giant["ring_params"] = [2.3, 0.7, 0.0, 0.0]
    giant["ring_color"] = [0.72, 0.68, 0.6, 0.0]
    x_axis = np.array([1.0, 0.0, 0.0])
    giant["axis_x"] = [*x_axis, 0.0]
    giant["axis_y"] = [*np.cross(pole, x_axis), 0.0]
    place(giant, direction(10.0, 18.0) * metres(4.5e8))
    moons = []
    for k in range(6):
        radius = metres(math.exp(rng.uniform(math.log(6e5), math.log(2.5e6))))
        albedo = [rng.uniform(0.1, 0.8) for _ in range(3)]
        m = body([0, 0, 0], radius, albedo, False)
        m["center_low"][3] = 0.6
        m["axis_x"] = [1, 0, 0, 0.0]
        m["axis_y"] = [0, 1, 0, float(k + 1)]
        az, el = rng.uniform(-35.0, 35.0), rng.uniform(8.0, 28.0)
        moons.append(place(m, direction(az, el) * metres(rng.uniform(2.0e7, 6.0e7))))
    sun = star(direction(-55.0, 20.0) * metres(1.5e11), metres(7.0e8))
    f.bodies = bodies(sun, host, giant, *moons)
    au = METRES_PER_AU
    g = f.globals
    g["atmo_center"] = [*split(host_c)[0], r]
    g["atmo_rayleigh"] = [5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au, 2.0e-5 * au]
    g["atmo_params"] = [0.76, 8000.0 / au, 80000.0 / au, 2.0]
    return f


def exponential_density(frame: Frame, bins: int = 16) -> np.ndarray:
    """exp(-h / H) at the middle altitude of each of ``bins`` slices."""
    top, scale = float(frame.globals["atmo_params"][2]), float(frame.globals["atmo_params"][1])
    return np.exp(-(np.arange(bins) + 0.5) * top / bins / scale).astype(np.float32)


def ring_profile_cba(frame: Frame, ring_body: int, bins: int = 512, seed: int = 0x5EED1234) -> np.ndarray:
    """A 512-bin ring optical-depth profile of C, B and A bands (0.1, 1.5,
    0.5) with 0.01 gaps (Cassini division and two narrow gaps in A) and
    +-10% noise from a PCG seeded by ``seed``, over the body's ring radii."""
    b = frame.bodies[ring_body]
    inner, outer = float(b["ring_plane"][3]), float(b["ring_params"][0])
    radius = inner + (np.arange(bins) + 0.5) / bins * (outer - inner)  # in planet radii
    # Saturn-like band edges scaled onto [inner, outer].
    edges = np.array([1.24, 1.53, 1.95, 2.03, 2.27]) - 1.24
    x = (radius - inner) / (outer - inner) * edges[-1]
    tau = np.where(x < edges[1], 0.1, np.where(x < edges[2], 1.5, np.where(x < edges[3], 0.01, 0.5)))
    for gap in (2.14 - 1.24, 2.21 - 1.24):
        tau = np.where(np.abs(x - gap) < 0.004, 0.01, tau)
    rng = Pcg(seed)
    noise = np.array([rng.uniform(-0.1, 0.1) for _ in range(bins)])
    return (tau * (1.0 + noise)).astype(np.float32)


DIFF_ALBEDO_BODIES = tuple(range(1, 9))
DIFF_RING_BODY = 2


def diff_system_parameters(frame: Frame, seed: int = 2, spread: float = 0.1) -> tuple[dict, dict]:
    """The 557 parameters of the differentiable workload (``diff_system``):
    the RGB albedos of its eight lit bodies (24), the C/B/A ring profile of
    the giant (512 bins), Rayleigh rgb, Mie and Mie g (5, per AU) and the
    16-slice exp(-h / H) density table. Returns the nominal set and a target
    set with every parameter scaled by its own factor in [1 - spread,
    1 + spread] from a PCG seeded by ``seed``. Each set is a dict of float32
    arrays: albedo (8, 3), tau (512,), air (5,), density (16,)."""
    g = frame.globals
    nominal = {
        "albedo": np.stack([frame.bodies["color"][b][:3] for b in DIFF_ALBEDO_BODIES]).astype(np.float32),
        "tau": ring_profile_cba(frame, DIFF_RING_BODY),
        "air": np.array([*g["atmo_rayleigh"], g["atmo_params"][0]], np.float32),
        "density": exponential_density(frame),
    }
    rng = Pcg(seed)
    target = {}
    for name, value in nominal.items():
        factors = np.array([1.0 + rng.uniform(-spread, spread) for _ in range(value.size)], np.float64)
        target[name] = (value.astype(np.float64) * factors.reshape(value.shape)).astype(np.float32)
    return nominal, target


def random_system(width: int, height: int, rng: "Pcg") -> tuple[Frame, dict]:
    """One random view of the batched workload: the camera 2 m above a host
    planet at a random pose (the host's up vector within 60 deg of the
    view's up, random heading), a star and 0-10 further bodies (moons and
    giants) at random directions, 2-12 bodies in all. A giant carries a
    ring half of the time (banded half of those), the host has Earth-like
    air a third of the time. Albedos are constant (style 0) or gas-giant
    bands (style 1) for giants."""
    f = frame()
    fov = rng.uniform(40.0, 75.0)
    camera(f.globals, fov, width, height)
    tilt = math.radians(rng.uniform(0.0, 60.0))
    heading = rng.uniform(0.0, 2.0 * math.pi)
    up = np.array([math.sin(tilt) * math.cos(heading), math.cos(tilt), math.sin(tilt) * math.sin(heading)])
    r_host = metres(math.exp(rng.uniform(math.log(1.5e6), math.log(7.0e6))))
    host_c = -up * (r_host + metres(2.0))
    host = place(body([0, 0, 0], r_host, [rng.uniform(0.1, 0.5) for _ in range(3)], False), host_c)
    sun = star(direction(rng.uniform(-80.0, 80.0), rng.uniform(5.0, 60.0)) * metres(1.5e11), metres(7.0e8))
    extra = []
    n = int(rng.uniform(0.0, 11.0))
    for k in range(n):
        giant = rng.uniform() < 0.3
        radius = metres(rng.uniform(2.0e7, 7.0e7) if giant else math.exp(rng.uniform(math.log(2e5), math.log(3e6))))
        dist = radius * rng.uniform(6.0, 40.0)
        b = body([0, 0, 0], radius, [rng.uniform(0.05, 0.8) for _ in range(3)], False)
        if giant:
            b["axis_x"] = [1, 0, 0, 1.0 if rng.uniform() < 0.5 else 0.0]
            b["axis_y"] = [0, 1, 0, float(k)]
            b["surface_color"] = [0.2, 0.17, 0.13, 0.8]
            b["surface_params"] = [7.0, 2.0, 0.0, 0.0]
            if rng.uniform() < 0.5:
                pole = direction(rng.uniform(-180.0, 180.0), rng.uniform(20.0, 90.0))
                b["ring_plane"] = [*pole, 1.2]
                b["ring_params"] = [2.3, rng.uniform(0.2, 1.2), 1.0 if rng.uniform() < 0.5 else 0.0, 0.0]
                b["ring_color"] = [0.7, 0.66, 0.6, 0.0]
        extra.append(place(b, direction(rng.uniform(-60.0, 60.0), rng.uniform(2.0, 50.0)) * dist))
    f.bodies = bodies(sun, host, *extra)
    air = rng.uniform() < 1.0 / 3.0
    if air:
        au = METRES_PER_AU
        g = f.globals
        g["atmo_center"] = [*split(host_c)[0], r_host]
        g["atmo_rayleigh"] = [5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au, 2.0e-5 * au]
        g["atmo_params"] = [0.76, 8000.0 / au, 80000.0 / au, 2.0]
    return f, dict(bodies=len(f.bodies), air=air)


def system_views(count: int = 96, seed: int = 1, min_pixels: int = 128 * 128, max_pixels: int = 1920 * 1080,
                 spp=(4, 8, 16, 32, 64), bounces: int = 4):
    """The batched-serving workload: ``count`` random systems
    (``random_system``) with pixel counts log-uniform in [min_pixels,
    max_pixels] at aspect 1, 4:3 or 16:9 and samples per pass from ``spp``,
    all from one PCG seeded by ``seed``."""
    rng = Pcg(seed)
    aspects = (1.0, 4.0 / 3.0, 16.0 / 9.0)
    views = []
    for _ in range(count):
        pixels = math.exp(rng.uniform(math.log(min_pixels), math.log(max_pixels)))
        aspect = aspects[min(2, int(rng.uniform(0.0, 3.0)))]
        height = max(8, int(round(math.sqrt(pixels / aspect))))
        width = max(8, int(round(height * aspect)))
        samples = int(spp[min(len(spp) - 1, int(rng.uniform(0.0, len(spp))))])
        f, info = random_system(width, height, rng)
        views.append(dict(kind="air" if info["air"] else "airless", frame=f, terrain=None, width=width,
                          height=height, spp=samples, bounces=bounces, bodies=info["bodies"]))
    return views
