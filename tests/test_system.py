"""Scene files, orbits and the observer (stargaze_amd.system) against the
reference's own assertions on the same bundled files (src/config.rs:661-1220,
src/sim.rs:860-1219), plus the frame packing against what the geometry fixes.
The scene files are the bundled copies in stargaze_amd/configs/."""
from __future__ import annotations

import math
import pathlib

import numpy as np
import pytest
import yaml

from stargaze_amd import ground as gr
from stargaze_amd import system as sy
from stargaze_amd.layouts import AU_KM, METRES_PER_AU

CONFIGS = pathlib.Path(sy.__file__).resolve().parent / "configs"


def text(name: str) -> str:
    return (CONFIGS / name).read_text(encoding="utf-8")


HALO, SOLAR, CORONA, UMBRA = (text(n) for n in ("halo.yaml", "solar-system.yaml", "corona-valley.yaml",
                                                 "3-umbra-valley.yaml"))


def by_name(scene: sy.Scene, name: str) -> int:
    return next(i for i, b in enumerate(scene.bodies) if b.name == name)


def angle(a: np.ndarray, b: np.ndarray) -> float:
    c = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
    return math.acos(min(1.0, max(-1.0, float(c))))


def startup(scene: sy.Scene) -> tuple[sy.Observer, float]:
    t = scene.initial_time_days
    return sy.Observer.for_scene(scene, t), t


def rejected(change, expected: str) -> None:
    """src/config.rs:1021-1032: a change to the halo definition is refused with
    a message naming the problem."""
    system = yaml.safe_load(HALO)
    change(system)
    with pytest.raises(sy.SystemError, match=expected):
        sy.parse(yaml.safe_dump(system))


# -- config.rs ----------------------------------------------------------------
def test_every_bundled_scene_loads():
    for path in sorted(CONFIGS.glob("*.yaml")):
        scene = sy.load(path)
        assert 2 <= len(scene.bodies) <= 256
        for i, b in enumerate(scene.bodies):
            assert b.parent is None or b.parent < i  # parents precede children


def test_bundled_scenes_and_unordered_bodies_load():
    """src/config.rs:1034-1062."""
    solar = sy.parse(SOLAR)
    assert len(solar.bodies) == 19 and solar.bodies[solar.host].name == "Earth"
    original = sy.parse(HALO)
    system = yaml.safe_load(HALO)
    system["bodies"].reverse()
    reordered = sy.parse(yaml.safe_dump(system))
    for t in (0.0, 12.3, -15.0, 10_000.0):
        before, after = original.positions(t), reordered.positions(t)
        for i, b in enumerate(original.bodies):
            assert np.linalg.norm(before[i] - after[by_name(reordered, b.name)]) < 1e-12
    halo, calyx = original.bodies[original.host], original.bodies[original.default_target]
    assert abs(halo.elements.a * AU_KM - 29919.57414) < 1e-6
    assert (halo.elements.a - halo.radius - calyx.radius * calyx.rings.outer_radius) * AU_KM > 10_000.0


def test_bundled_levels_have_atmospheres_and_airless_configs_still_load():
    """src/config.rs:1064-1082."""
    for name in ("halo.yaml", "solar-system.yaml", "puzzle.yaml", "median-resonance.yaml", "vesper.yaml"):
        assert sy.parse(text(name)).atmosphere is not None
    dense = sy.parse(text("median-resonance.yaml")).atmosphere
    assert dense.mie > sy.Atmosphere.earthlike().mie
    assert sy.parse(HALO.replace("  atmosphere: earthlike\n", "")).atmosphere is None
    with pytest.raises(sy.SystemError):
        sy.parse(HALO.replace("earthlike", "unknown"))


def test_target_lists_can_span_multiple_pages():
    """src/config.rs:1084-1103."""
    system = yaml.safe_load(SOLAR)
    system["targets"] = [b["id"] for b in system["bodies"]
                         if b["id"] != system["camera"]["body"] and b["kind"] != "anchor"]
    count = len(system["targets"])
    assert count > 10
    scene = sy.parse(yaml.safe_dump(system))
    assert len(scene.targets) == count
    assert all(i < len(scene.bodies) and i != scene.host for i in scene.targets)


def test_validates_optional_camera_direction():
    """src/config.rs:1105-1136."""
    assert sy.parse(HALO).initial_direction_deg is None
    assert sy.parse(text("vesper.yaml")).initial_direction_deg == (0.0, 25.0)
    for azimuth in (-1.0, 360.0, float("nan")):
        rejected(lambda s: s["camera"].update(direction={"azimuth_deg": azimuth, "altitude_deg": 25.0}),
                 "direction.azimuth_deg")
    for altitude in (-91.0, 91.0, float("inf")):
        rejected(lambda s: s["camera"].update(direction={"azimuth_deg": 0.0, "altitude_deg": altitude}),
                 "direction.altitude_deg")


def test_rejects_bad_structure_and_references():
    """src/config.rs:1138-1161."""
    rejected(lambda s: s.update(version=2), "unsupported version")
    rejected(lambda s: s["bodies"][1].update(id=s["bodies"][0]["id"]), "duplicate body")
    rejected(lambda s: s["bodies"][6].update(parent="missing"), "unknown parent")
    rejected(lambda s: s["bodies"][5].update(parent="halo"), "cyclic parent")
    rejected(lambda s: s["camera"].update(body="missing"), "camera.body")
    rejected(lambda s: s["camera"].update(look_at=s["camera"]["body"]), "other than the camera host")
    rejected(lambda s: s.setdefault("targets", []).append("missing"), "unknown body")
    with pytest.raises(sy.SystemError):
        sy.parse(HALO + "\nmisspelled_setting: true\n")


def test_rejects_unsafe_numeric_values():
    """src/config.rs:1163-1190."""
    orbit = lambda s: s["bodies"][6]["orbit"]  # noqa: E731
    rejected(lambda s: orbit(s).update(eccentricity=1.0), "eccentricity")
    rejected(lambda s: orbit(s).update(period_days=0.0), "orbital period")
    rejected(lambda s: orbit(s).update(semi_major_axis_km=100.0), "intersects the parent")
    rejected(lambda s: orbit(s).update(semi_major_axis_au=1.0), "exactly one")
    rejected(lambda s: s["bodies"][6].update(radius_km=float("nan")), "radius_km")
    rejected(lambda s: s["bodies"][5]["rings"].update(outer_radius=1.0), "ring radii")
    rejected(lambda s: s["bodies"][5]["albedo"].__setitem__(0, 2.0), "albedo")
    rejected(lambda s: s["camera"].update(latitude_deg=91.0), "latitude_deg")
    rejected(lambda s: s["camera"].update(fov_deg=float("inf")), "fov_deg")
    rejected(lambda s: s["camera"].update(height_m=0.0), "height_m")
    rejected(lambda s: s.update(time_days=float("nan")), "time_days")


def test_tidal_rotation_aligns_with_arbitrary_orbit_and_epoch():
    """src/config.rs:1192-1219: longitude 180 of a locked moon faces its
    parent at any orbit orientation and epoch, and the eye stands at the
    configured height above the surface."""
    system = yaml.safe_load(HALO)
    system["bodies"][6]["orbit"].update(inclination_deg=47.0, ascending_node_deg=123.0, periapsis_deg=71.0,
                                        mean_anomaly_deg=23.0, epoch_days=52.0)
    system["time_days"], system["camera"]["height_m"] = -42.0, 25.0
    scene = sy.parse(yaml.safe_dump(system))
    moon = scene.bodies[scene.host]
    for t in (-100.0, 0.0, 52.0, 1000.0):
        facing = moon.spin(t) @ np.array([-1.0, 0.0, 0.0])
        rel = moon.relative_position(t)
        assert facing @ (-rel / np.linalg.norm(rel)) > 1.0 - 1e-12
    observer, t = startup(scene)
    assert t == -42.0 and observer.height_m == 25.0
    positions = scene.positions(t)
    vf = observer.frame(scene, t, positions)
    altitude = (np.linalg.norm(vf.position - positions[scene.host]) - moon.radius) * AU_KM * 1000.0
    assert abs(altitude - 25.0) < 0.001


def test_gas_giant_surface_defaults_and_validation():
    """src/config.rs:766-797."""
    scene = sy.parse(CORONA)
    tormenta = scene.bodies[by_name(scene, "Tormenta")]
    s = tormenta.surface
    assert s.style == sy.STYLE_CODES["gas_giant"] and (s.seed, s.storms) == (21, 3)
    zone, belt = tormenta.albedo, s.secondary_albedo
    assert all(belt[c] < zone[c] for c in range(3))
    assert belt[2] / belt[0] < zone[2] / zone[0]
    line = "surface: {style: gas_giant, seed: 21, bands: 8, turbulence: 0.7, storms: 3}"
    assert line in CORONA
    for bad in ("surface: {style: gas_giant, bands: 0}", "surface: {style: gas_giant, turbulence: 2}",
                "surface: {style: gas_giant, storms: 4}", "surface: {style: gas_giant, belt_albedo: [2, 0, 0]}",
                "surface: {style: marble}", "surface: {style: gas_giant, colour: 1}",
                "surface: {style: gas_giant, craters: 0.5}"):
        with pytest.raises(sy.SystemError):
            sy.parse(CORONA.replace(line, bad))
    rotation = "    rotation: {mode: free, period_days: 9, obliquity_deg: 25}"
    star = CORONA.replace(rotation, rotation + "\n    surface: {style: gas_giant}", 1)
    with pytest.raises(sy.SystemError):
        sy.parse(star)


def test_solid_surface_styles_defaults_and_field_checks():
    """src/config.rs:799-844."""
    scene = sy.parse(CORONA)
    surface = lambda name: scene.bodies[by_name(scene, name)].surface  # noqa: E731
    ocre = surface("Ocre")
    assert ocre.style == sy.STYLE_CODES["rocky"] and ocre.craters > 0.0 and ocre.ice_caps > 0.0
    assert surface("Chispa").style == sy.STYLE_CODES["cratered"]
    assert surface("Ceniza").style == sy.STYLE_CODES["volcanic"]
    assert surface("Alba").style == sy.STYLE_CODES["icy"]
    puna = surface("Puna")
    assert puna.style == sy.STYLE_CODES["terran"] and puna.ocean > 0.5 and puna.clouds > 0.0
    assert scene.bodies[scene.host].surface is None
    line = "surface: {style: rocky, seed: 1}"
    assert line in CORONA
    sy.parse(CORONA.replace(line, "surface: {style: rocky, craters: 1, ice_caps: 0}"))
    for bad in ("surface: {style: rocky, ocean: 0.5}", "surface: {style: rocky, clouds: 0.5}",
                "surface: {style: rocky, bands: 4}", "surface: {style: rocky, storms: 1}",
                "surface: {style: rocky, craters: 1.5}", "surface: {style: rocky, ice_caps: -0.1}"):
        with pytest.raises(sy.SystemError):
            sy.parse(CORONA.replace(line, bad))
    terran = "surface: {style: terran, seed: 7}"
    sy.parse(CORONA.replace(terran, "surface: {style: terran, ocean: 0.3, clouds: 0}"))
    with pytest.raises(sy.SystemError):
        sy.parse(CORONA.replace(terran, "surface: {style: terran, ocean: 2}"))


def test_default_secondary_albedos_follow_the_style_rules():
    """The defaults of SurfaceDef::build (src/config.rs:217-229) restated per
    style on one albedo, checked by their defining properties: the same hue
    direction, darker, and the fixed tints of icy, volcanic and terran."""
    albedo = (0.8, 0.6, 0.4)
    for style in sy.STYLE_CODES:
        body = {"id": "b", "kind": "planet", "parent": "s", "radius_km": 1000.0, "albedo": list(albedo),
                "orbit": {"semi_major_axis_au": 1.0, "period_days": 365.0}, "surface": {"style": style}}
        star = {"id": "s", "kind": "star", "radius_km": 7e5, "luminosity": 1.0}
        system = {"version": 1, "name": "x", "bodies": [star, body],
                  "camera": {"body": "b", "latitude_deg": 0.0, "longitude_deg": 0.0, "look_at": "s",
                             "fov_deg": 10.0}}
        second = np.array(sy.parse(yaml.safe_dump(system)).bodies[1].surface.secondary_albedo)
        a = np.array(albedo, np.float32)
        expect = {
            "gas_giant": 0.8 * a * (a / a.max()) ** 1.6, "cratered": 0.55 * a * (a / a.max()) ** 0.6,
            "rocky": 0.6 * a * (a / a.max()) ** 1.0, "icy": a * [0.75, 0.5, 0.35],
            "volcanic": a * [0.85, 0.45, 0.2], "terran": np.array([0.03, 0.09, 0.2]),
        }[style]
        np.testing.assert_allclose(second, expect, rtol=1e-6)


def test_gemma_and_gemmi_are_mutually_tidally_locked():
    """src/config.rs:846-866."""
    erg = text("2-erg-dunes.yaml")
    scene = sy.parse(erg)
    gemma, gemmi = by_name(scene, "Gemma"), by_name(scene, "Gemmi")
    x = np.array([1.0, 0.0, 0.0])
    for t in (0.0, 0.7, 1.6, 2.9, 40.0):
        p = scene.positions(t)
        towards = (p[gemmi] - p[gemma]) / np.linalg.norm(p[gemmi] - p[gemma])
        assert (scene.bodies[gemma].spin(t) @ x) @ towards > 0.9999
        assert (scene.bodies[gemmi].spin(t) @ x) @ towards > 0.9999
    lock = "rotation: {mode: tidal_lock, to: gemmi}"
    assert lock in erg
    for bad in ("rotation: {mode: tidal_lock, to: nowhere}", "rotation: {mode: tidal_lock, to: simoom}"):
        with pytest.raises(sy.SystemError):
            sy.parse(erg.replace(lock, bad))


def test_umbra_host_is_locked_with_two_moons_and_daily_giant_eclipses():
    """src/config.rs:668-739."""
    scene = sy.parse(UMBRA)
    star, giant = by_name(scene, "Lucero"), by_name(scene, "Umbra")
    host = scene.bodies[scene.host]
    assert host.tidally_locked and scene.ground == sy.THEMES["alpine"] and scene.day_length_days == 1.0
    s, g = scene.bodies[star], scene.bodies[giant]
    assert s.parent == g.parent and s.period_days == g.period_days
    assert g.radius > 10.0 * s.radius and g.rings is not None and s.rings is None
    assert sum(1 for b in scene.bodies if b.parent == scene.host and b.kind == sy.KIND_MOON) == 2
    inner = [b for b in scene.bodies if b.kind == sy.KIND_PLANET and b.parent == host.parent
             and g.elements.a < b.elements.a < host.elements.a]
    assert len(inner) == 2 and all(b.radius < host.radius and b.elements.inc == 0.0 for b in inner)
    observer, _ = startup(scene)
    for t in (0.0, 1.0, 2.0, 19.0, 50.0):
        p = scene.positions(t)
        vf = observer.frame(scene, t, p)
        ds, dg = p[star] - vf.position, p[giant] - vf.position
        rs, rg = math.asin(s.radius / np.linalg.norm(ds)), math.asin(g.radius / np.linalg.norm(dg))
        assert np.linalg.norm(dg) < np.linalg.norm(ds)
        assert angle(ds, dg) + rs < rg, f"no totality on day {t}"
        altitude = math.degrees(math.asin(ds @ vf.zenith / np.linalg.norm(ds)))
        assert 49.0 < altitude < 51.0, f"star altitude {altitude} on day {t}"
    for duration in ("0", "-1", ".inf", ".nan"):
        with pytest.raises(sy.SystemError):
            sy.parse(UMBRA.replace("day_length_days: 1", f"day_length_days: {duration}"))
    assert sy.parse(UMBRA.replace("day_length_days: 1\n", "")).day_length_days is None


def test_level_one_opens_by_day_and_calyx_hides_both_suns_before_they_set():
    """src/config.rs:868-906."""
    scene = sy.parse(text("1-halo-volcano.yaml"))
    observer, start = startup(scene)
    calyx, suns = by_name(scene, "Calyx"), [by_name(scene, "Aur"), by_name(scene, "Igni")]

    def sky(t):
        p = scene.positions(t)
        vf = observer.frame(scene, t, p)
        dc = p[calyx] - vf.position
        rc = math.asin(scene.bodies[calyx].radius / np.linalg.norm(dc))
        out = []
        for s in suns:
            ds = p[s] - vf.position
            rs = math.asin(scene.bodies[s].radius / np.linalg.norm(ds))
            alt = math.degrees(math.asin(ds @ vf.zenith / np.linalg.norm(ds)))
            out.append((alt, angle(ds, dc) < rc - rs))
        return out

    assert all(alt > 25.0 and not hidden for alt, hidden in sky(start))
    times = [start + i * 0.005 for i in range(2000)]
    both = next(t for t in times if all(hidden for _, hidden in sky(t)))
    set_ = next(t for t in times if all(alt < 0.0 for alt, _ in sky(t)))
    assert both - start > 1.0
    assert all(alt > 5.0 for alt, _ in sky(both))
    assert set_ - both > 0.5


def test_corona_level_opens_before_puna_occults_the_ringed_star_at_night():
    """src/config.rs:908-935."""
    scene = sy.parse(CORONA)
    assert scene.ground == sy.THEMES["desert"]
    corona, brasa, puna = by_name(scene, "Corona"), by_name(scene, "Brasa"), by_name(scene, "Puna")
    assert scene.bodies[corona].kind == sy.KIND_STAR and scene.bodies[corona].rings is not None
    assert not scene.bodies[scene.host].tidally_locked
    assert scene.bodies[scene.host].parent == brasa and scene.bodies[puna].parent == brasa
    observer, start = startup(scene)

    def total(t):
        p = scene.positions(t)
        vf = observer.frame(scene, t, p)
        dc, dp = p[corona] - vf.position, p[puna] - vf.position
        rc = math.asin(scene.bodies[corona].radius / np.linalg.norm(dc))
        rp = math.asin(scene.bodies[puna].radius / np.linalg.norm(dp))
        up = lambda d: math.degrees(math.asin(d @ vf.zenith / np.linalg.norm(d)))  # noqa: E731
        return angle(dc, dp) < rp - rc and up(dc) > 20.0 and up(p[brasa] - vf.position) < -10.0

    assert any(total(start + i * 0.00005) for i in range(2400)), f"no total occultation after day {start}"


def test_desert_level_has_free_moon_submoon_and_visible_wanderers():
    """src/config.rs:937-1016, including the frame's ground of tumbled rocks."""
    scene = sy.parse(text("atacama.yaml"))
    assert scene.ground == sy.THEMES["desert"]
    assert sy.parse(HALO).ground == sy.THEMES["forest"]
    with pytest.raises(sy.SystemError):
        sy.parse(HALO.replace("  atmosphere: earthlike", "  ground: unknown\n  atmosphere: earthlike"))
    host = scene.bodies[scene.host]
    assert not host.tidally_locked and host.spin_period != host.period_days
    giant = host.parent
    moons = sorted(((i, b) for i, b in enumerate(scene.bodies) if b.parent == giant), key=lambda m: m[1].elements.a)
    assert len(moons) == 3 and moons[1][0] == scene.host
    ring = scene.bodies[giant].rings
    assert ring.outer_radius < 1.5
    assert ring.outer_radius * scene.bodies[giant].radius < moons[0][1].elements.a - moons[0][1].radius
    submoons = [b for b in scene.bodies if b.parent == scene.host]
    assert len(submoons) == 1 and submoons[0].radius < host.radius * 0.1
    others = [(i, b) for i, b in enumerate(scene.bodies) if i != giant and b.kind == sy.KIND_PLANET]
    assert len(others) == 3
    assert sum(1 for _, b in others if b.elements.a < scene.bodies[giant].elements.a) == 1
    observer, t = startup(scene)
    positions = scene.positions(t)
    view = observer.frame(scene, t, positions)
    assert 0.2 < view.sun_altitude < 0.4
    for i, b in others:
        d = positions[i] - view.position
        assert d @ view.zenith / np.linalg.norm(d) > 0.2, f"{b.name} below the horizon"
        assert b.luminosity == 0.0
    frame = sy.build_frame(scene, t, 960, 540, observer=observer)
    assert frame.globals["ground_counts"][2] == 1
    assert np.all((frame.ground["center"][:, 3] < 0.0) | (frame.ground["center"][:, 3] == gr.TUMBLED_ROCK))


# -- sim.rs -------------------------------------------------------------------
def test_earthlike_atmosphere_uses_consistent_units():
    """src/sim.rs:861-872."""
    a = sy.Atmosphere.earthlike()
    assert abs(a.scale_height * 149_597_870_700.0 - 8_500.0) < 1e-9
    assert abs(a.thickness * 149_597_870_700.0 - 80_000.0) < 1e-9
    assert abs((a.rayleigh[2] + a.mie) * a.scale_height - 0.45985) < 1e-6


def test_own_hue_and_unit_density_reproduce_the_air():
    """src/sim.rs:1200-1219."""
    air = sy.Atmosphere.earthlike()
    same = air.tuned(1.0, air.hue())
    for a, b in zip(air.rayleigh, same.rayleigh):
        assert abs(a - b) <= 1e-5 * a
    assert 200.0 <= air.hue() < 240.0
    red = air.tuned(2.0, 0.0)
    assert red.rayleigh[0] > red.rayleigh[2] and red.mie == air.mie * 2.0


def test_hsv_round_trip_and_primaries():
    """hsv and rgb invert each other; the primaries sit at 0, 120 and 240 deg."""
    for c, hue in (((1.0, 0.0, 0.0), 0.0), ((0.0, 1.0, 0.0), 120.0), ((0.0, 0.0, 1.0), 240.0)):
        assert sy.hsv(np.array(c))[0] == hue
    rng = np.random.default_rng(5)
    for c in rng.uniform(0.0, 1.0, (200, 3)).astype(np.float32):
        np.testing.assert_allclose(sy.rgb(*sy.hsv(c)), c, atol=2e-6)


def test_solar_system_has_real_scales_and_all_observable_planets():
    """src/sim.rs:883-937 on the bundled solar system file."""
    scene = sy.parse(SOLAR)
    assert scene.bodies[scene.host].name == "Earth"
    assert sum(b.kind == sy.KIND_PLANET for b in scene.bodies) == 8
    assert sum(b.kind == sy.KIND_MOON for b in scene.bodies) == 10
    for name, radius_km, distance_au, period_days in (
        ("Mercury", 2_439.7, 0.38709927, 87.9691), ("Venus", 6_051.8, 0.72333566, 224.701),
        ("Earth", 6_371.0, 1.00000261, 365.256), ("Mars", 3_389.5, 1.52371034, 686.980),
        ("Jupiter", 69_911.0, 5.202887, 4_332.589), ("Saturn", 58_232.0, 9.53667594, 10_759.22),
        ("Uranus", 25_362.0, 19.18916464, 30_688.5), ("Neptune", 24_622.0, 30.06992276, 60_182.0),
    ):
        i = by_name(scene, name)
        b = scene.bodies[i]
        assert abs(b.radius * AU_KM - radius_km) < 1e-6
        assert abs(b.elements.a - distance_au) < 1e-9
        assert abs(b.period_days - period_days) < 1e-6
        assert (i in scene.targets) == (i != scene.host)
        assert np.linalg.norm(b.relative_position(0.0) - b.relative_position(b.period_days)) < 1e-10
    assert len(scene.targets) == 9
    assert abs(scene.bodies[by_name(scene, "Phobos")].elements.a * AU_KM - 9_376.0) < 1e-6
    for t in (-365.25, 0.0, 365.25, 60_182.0):
        assert np.all(np.isfinite(scene.positions(t)))
    for moon in (b for b in scene.bodies if b.kind == sy.KIND_MOON):  # src/sim.rs:874-881
        parent = scene.bodies[moon.parent]
        assert moon.elements.a * (1.0 - moon.elements.e) > parent.radius + moon.radius


def test_spin_keeps_latitude_fixed_and_respects_axial_tilt():
    """src/sim.rs:939-968."""
    scene = sy.parse(SOLAR)
    lat = math.radians(35.0)
    local = np.array([math.cos(lat), 0.0, math.sin(lat)])
    z = np.array([0.0, 0.0, 1.0])
    for b in scene.bodies:
        if b.tidally_locked:
            continue
        pole = sy.rot_y(b.obliquity) @ z
        for fraction in (0.0, 0.25, 0.5, 0.75, 1.0):
            q = b.spin(b.spin_period * fraction)
            assert np.linalg.norm(q @ z - pole) < 1e-12
            assert abs((q @ local) @ pole - math.sin(lat)) < 1e-12
    for name in ("Venus", "Uranus"):
        assert (scene.bodies[by_name(scene, name)].spin(0.0) @ z)[2] < 0.0
    for name, inclination in (("Phobos", 1.075), ("Titania", 0.079), ("Triton", 156.865)):
        moon = scene.bodies[by_name(scene, name)]
        pole = scene.bodies[moon.parent].spin(0.0) @ z
        normal = sy.rot_z(moon.elements.node) @ sy.rot_x(moon.elements.inc) @ z
        assert abs(pole @ normal - math.cos(math.radians(inclination))) < 1e-12


def test_mars_closes_its_orbit_only_after_its_own_period():
    """src/sim.rs:1085-1114."""
    scene = sy.parse(SOLAR)
    mars, earth = by_name(scene, "Mars"), by_name(scene, "Earth")
    assert abs(scene.bodies[earth].period_days - 365.256) < 0.01
    assert abs(scene.bodies[mars].period_days - 686.98) < 0.5
    start = scene.positions(0.0)[mars]
    assert np.linalg.norm(scene.positions(686.980)[mars] - start) < 0.02
    assert np.linalg.norm(scene.positions(365.256)[mars] - start) > 0.3


def test_kepler_solution_satisfies_the_equation_and_the_ellipse():
    """E - e sin E = M to rounding, and the position lies on the ellipse with
    the focus at the parent: |r| = a (1 - e^2) / (1 + e cos nu)."""
    for e in (0.0, 0.2, 0.7, 0.85, 0.97):
        for m in np.linspace(-7.0, 7.0, 29):
            ea = sy.solve_kepler(m, e)
            assert abs(ea - e * math.sin(ea) - m % (2 * math.pi)) < 1e-12
        b = sy.Body("x", sy.KIND_MOON, 0, sy.Elements(2.0, e, 0.3, 1.1, 0.7, 0.2, 5.0), 10.0, 0.01, (0.5,) * 3,
                    (1.0,) * 3, 0.0, 0.0, 1.0, 0.0)
        for t in np.linspace(0.0, 10.0, 17):
            r = b.relative_position(t)
            normal = sy.orbit_rotation(b.elements) @ np.array([0.0, 0.0, 1.0])
            assert abs(r @ normal) < 1e-12
            periapsis = sy.orbit_rotation(b.elements) @ np.array([1.0, 0.0, 0.0])
            nu = math.atan2(r @ np.cross(normal, periapsis), r @ periapsis)
            assert abs(np.linalg.norm(r) - 2.0 * (1 - e * e) / (1 + e * math.cos(nu))) < 1e-12


# -- the frame (main.rs) ------------------------------------------------------
def test_frame_packs_bodies_in_telescope_space():
    """build_frame on the level-one file: each body's f32 high and low centre
    parts sum to the f64 telescope-space vector, stars carry 2.5 L / r^2 in
    their emission, ring poles and spin axes are unit vectors, anchors are
    left out and the air shell points at the host."""
    scene = sy.parse(text("1-halo-volcano.yaml"))
    observer, t = startup(scene)
    frame = sy.build_frame(scene, t, 320, 180, observer=observer)
    positions = scene.positions(t)
    vf = observer.frame(scene, t, positions)
    visible = [i for i, b in enumerate(scene.bodies) if b.kind != sy.KIND_ANCHOR]
    assert len(frame.bodies) == len(visible)
    for rec, i in zip(frame.bodies, visible):
        b = scene.bodies[i]
        view = vf.world_to_view(positions[i] - vf.position)
        summed = rec["center"].astype(np.float64) + rec["center_low"][:3].astype(np.float64)
        assert np.linalg.norm(summed - view) <= 1e-13 * np.linalg.norm(view) + 1e-20
        assert rec["radius"] == np.float32(b.radius)
        if b.kind == sy.KIND_STAR:
            radiance = np.float32(2.5 * b.luminosity / b.radius ** 2)
            np.testing.assert_array_equal(rec["color"], np.float32(b.emission) * radiance)
            assert rec["emissive"] == 1.0 and rec["center_low"][3] == 0.0
        else:
            np.testing.assert_array_equal(rec["color"], np.float32(b.albedo))
        for axis in ("axis_x", "axis_y"):
            assert abs(np.linalg.norm(rec[axis][:3]) - 1.0) < 1e-6
        if b.rings is not None:
            assert abs(np.linalg.norm(rec["ring_plane"][:3]) - 1.0) < 1e-6
            assert rec["ring_plane"][3] == np.float32(b.rings.inner_radius)
    g = frame.globals
    host_index = visible.index(scene.host)
    assert g["atmo_params"][3] == host_index + 1 and g["ground_north"][3] == host_index + 1
    host_view = vf.world_to_view(positions[scene.host] - vf.position)
    np.testing.assert_allclose(g["atmo_center"][:3], host_view, rtol=1e-6)
    # The eye is the configured height above the host: the host centre lies
    # straight down (ground_up points away from it) at radius + height.
    up = g["ground_up"][:3].astype(np.float64)
    depth = np.linalg.norm(host_view) - scene.bodies[scene.host].radius
    assert abs(depth * METRES_PER_AU - scene.observer_height_m) < 1e-3
    assert np.linalg.norm(host_view / np.linalg.norm(host_view) + up) < 1e-6
    assert g["ground_counts"][2] == sy.THEMES["basalt"] and g["ground_counts"][1] == 1
    assert g["viewport"][0] == 320 and g["viewport"][1] == 180
    assert abs(g["cam_forward"][3] - math.tan(math.radians(70.0) / 2)) < 1e-6


def test_point_at_fov_centres_the_target():
    """Without a stored direction the startup view aims at look_at: the
    target's telescope-space centre lies on the -Z axis."""
    scene = sy.parse(HALO)
    t = 17.25
    frame = sy.build_frame(scene, t, 64, 64)
    target = [i for i, b in enumerate(scene.bodies) if b.kind != sy.KIND_ANCHOR].index(scene.default_target)
    c = frame.bodies[target]["center"].astype(np.float64) + frame.bodies[target]["center_low"][:3]
    assert c[2] < 0.0 and math.hypot(c[0], c[1]) < 1e-9 * abs(c[2])
    assert abs(frame.globals["cam_forward"][3] - math.tan(math.radians(scene.default_fov_deg) / 2)) < 1e-7


def test_airless_ground_and_air_controls():
    """The crater ground removes the host's air, air density below 0.005
    removes it too, and a density scales Rayleigh and Mie
    (src/main.rs:915-925)."""
    halo = sy.parse(HALO)
    t = 3.0
    assert sy.build_frame(halo, t, 32, 32).globals["atmo_rayleigh"][3] > 0.0
    craters = sy.parse(HALO.replace("  atmosphere: earthlike", "  atmosphere: earthlike\n  ground: craters"))
    assert craters.ground == sy.THEMES["craters"]
    assert np.all(sy.build_frame(craters, t, 32, 32).globals["atmo_rayleigh"] == 0.0)
    assert sy.build_frame(craters, t, 32, 32, ground=False).globals["atmo_rayleigh"][3] > 0.0
    assert np.all(sy.build_frame(halo, t, 32, 32, air_density=0.004).globals["atmo_params"] == 0.0)
    one, two = (sy.build_frame(halo, t, 32, 32, air_density=d).globals["atmo_rayleigh"] for d in (1.0, 2.0))
    np.testing.assert_allclose(two, 2.0 * one, rtol=1e-6)
    no_ground = sy.build_frame(halo, t, 32, 32, ground=False)
    assert len(no_ground.ground) == 0 and no_ground.globals["ground_north"][3] == 0.0
    assert list(no_ground.globals["ground_counts"]) == [0, 0, 0, 0]


def test_terrain_lifts_the_eye_and_seats_the_scenery():
    """With a heightfield the startup eye stands its configured height above
    the terrain at the lookout, and rock bases follow the surface."""
    from stargaze_amd import terrain as tr

    scene = sy.parse(text("1-halo-volcano.yaml"))
    heights = tr.ridged_fbm(64, seed=0x743337A9, amplitude=60.0)
    radius_m = scene.bodies[scene.host].radius * METRES_PER_AU
    lay = sy.surface_height((heights, 64), radius_m)
    t = scene.initial_time_days
    flat = sy.build_frame(scene, t, 64, 36)
    lifted = sy.build_frame(scene, t, 64, 36, terrain=(heights, 64))
    assert abs(lifted.globals["ground_east"][3] - (scene.observer_height_m + lay(0.0, 0.0))) < 1e-4
    assert flat.globals["ground_east"][3] == np.float32(scene.observer_height_m)
    assert len(flat.ground) == len(lifted.ground) > 0
    moved = np.abs(lifted.ground["center"][:, 1] - flat.ground["center"][:, 1])
    assert moved.max() > 0.1


def test_unbuilt_basalt_ground_is_refused_by_the_oracle():
    """The level-one file's basalt ground packs (theme 4) but has no
    restated materials: the oracle refuses it as the library does."""
    import oracle
    from stargaze_amd import terrain as tr
    from stargaze_amd.pathtracer import Options

    scene = sy.parse(text("1-halo-volcano.yaml"))
    frame = sy.build_frame(scene, scene.initial_time_days, 16, 9)
    orc = oracle.OracleTracer((16, 9), Options(samples_per_frame=1, max_bounces=1, milky_way=0.0), "f32")
    orc.upload_terrain(tr.terrain_data(tr.ridged_fbm(16), 16), 16)
    with pytest.raises(ValueError, match="theme 4"):
        orc.encode(frame)
    scene.ground = sy.THEMES["desert"]
    orc.encode(sy.build_frame(scene, scene.initial_time_days, 16, 9))
    assert np.all(np.isfinite(orc.read_accumulation()))
