"""Frames from the reference's scene files: YAML system format v1, Kepler
orbits in float64, the observer's surface frame and the telescope-space
packing of ``Globals`` and ``Sphere`` records.

This restates the host side that feeds the reference's path tracer:
``config::parse`` (src/config.rs:12-686), the orbit and spin model of
``sim::Body`` (src/sim.rs:146-244), ``Atmosphere`` (src/sim.rs:246-310),
``Observer::frame`` (src/camera.rs:53-121), the aim of ``point_at_fov``
(src/main.rs:392-400) and ``build_frame_with_vegetation``
(src/main.rs:822-1018), so a file such as ``configs/legacy/halo.yaml`` renders
through ``PathTracer.encode`` directly.

The search for a night view when a file has no ``time_days`` is not
restated: callers pass a time. ``build_frame`` takes the ground's heightfield
in the ``terrain`` module's layout (``terrain.generate`` restates the
reference's), seats the scenery on it and lifts the eye onto it as
``terrain::lookout_height`` does (src/terrain.rs:64-71).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import ground as gr
from . import terrain as tr
from .layouts import AU_KM, GLOBALS, METRES_PER_AU, PRIMITIVE, SPHERE
from .pathtracer import Frame

SUN_LIGHT = 2.5  # src/main.rs:71
CONFIGS = Path(__file__).resolve().parent / "configs"  # the reference's bundled scene files
MIN_FOV_DEG, MAX_FOV_DEG = 0.001, 90.0

KIND_ANCHOR, KIND_STAR, KIND_PLANET, KIND_MOON = "anchor", "star", "planet", "moon"
STYLE_CODES = {"gas_giant": 1, "cratered": 2, "rocky": 3, "icy": 4, "volcanic": 5, "terran": 6}
THEMES = {"forest": 0, "desert": 1, "craters": 2, "alpine": 3, "basalt": 4, "dunes": 5}


class SystemError(ValueError):
    """An invalid scene file (the reference's ``anyhow`` errors)."""


def _ensure(condition, message: str) -> None:
    if not condition:
        raise SystemError(message)


# -- rotations (glam conventions: right-handed, column vectors) ---------------
def rot_x(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def rot_y(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def rot_z(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


@dataclass
class Elements:
    """Keplerian elements (src/sim.rs:22-45): AU, radians, days."""

    a: float
    e: float = 0.0
    inc: float = 0.0
    node: float = 0.0
    peri: float = 0.0
    m0: float = 0.0
    epoch: float = 0.0


@dataclass
class Rings:
    inner_radius: float
    outer_radius: float
    optical_depth: float
    albedo: tuple
    banded: bool = False


@dataclass
class Surface:
    """Procedural surface look (src/sim.rs:69-144)."""

    style: int
    seed: int
    secondary_albedo: tuple
    bands: float
    turbulence: float
    storms: int
    craters: float
    ice_caps: float
    ocean: float
    clouds: float


@dataclass
class Body:
    """``sim::Body`` (src/sim.rs:146-175)."""

    name: str
    kind: str
    parent: int | None
    elements: Elements
    period_days: float
    radius: float  # AU
    albedo: tuple
    emission: tuple
    luminosity: float
    obliquity: float
    spin_period: float
    spin_phase: float
    rings: Rings | None = None
    surface: Surface | None = None
    tidally_locked: bool = False
    lock_to: tuple | None = None  # (Elements, period) of the moon it is locked to

    def relative_position(self, t: float) -> np.ndarray:
        """Position relative to the parent, AU (src/sim.rs:177-194)."""
        n = 2.0 * math.pi / self.period_days
        m = self.elements.m0 + n * (t - self.elements.epoch)
        e = self.elements.e
        ea = solve_kepler(m, e)
        r = self.elements.a * (1.0 - e * math.cos(ea))
        half = ea * 0.5
        nu = 2.0 * math.atan2(math.sqrt(1.0 + e) * math.sin(half), math.sqrt(1.0 - e) * math.cos(half))
        in_plane = np.array([r * math.cos(nu), r * math.sin(nu), 0.0])
        return orbit_rotation(self.elements) @ in_plane

    def lock_orbit(self) -> tuple:
        return self.lock_to if self.lock_to is not None else (self.elements, self.period_days)

    def spin(self, t: float) -> np.ndarray:
        """Orientation of the body's fixed frame (src/sim.rs:210-228)."""
        if self.tidally_locked:
            elements, period = self.lock_orbit()
            phase = elements.m0 + 2.0 * math.pi * (t - elements.epoch) / period
            return orbit_rotation(elements) @ rot_z(phase)
        theta = self.spin_phase + 2.0 * math.pi * (t / self.spin_period)
        return rot_y(self.obliquity) @ rot_z(theta)


def orbit_rotation(el: Elements) -> np.ndarray:
    return rot_z(el.node) @ rot_x(el.inc) @ rot_z(el.peri)


def solve_kepler(m: float, e: float) -> float:
    """E - e sin E = M by Newton's method (src/sim.rs:231-244)."""
    m = m % (2.0 * math.pi)
    ea = m if e < 0.8 else math.pi
    for _ in range(16):
        f = ea - e * math.sin(ea) - m
        fp = 1.0 - e * math.cos(ea)
        step = f / fp
        ea -= step
        if abs(step) < 1e-13:
            break
    return ea


@dataclass
class Atmosphere:
    """Single-scattering air in AU units (src/sim.rs:246-310)."""

    scale_height: float
    thickness: float
    rayleigh: tuple
    mie: float
    mie_g: float

    @staticmethod
    def earthlike() -> "Atmosphere":
        au = METRES_PER_AU
        return Atmosphere(8_500.0 / au, 80_000.0 / au, (5.8e-6 * au, 13.5e-6 * au, 33.1e-6 * au), 21.0e-6 * au, 0.76)

    @staticmethod
    def dense() -> "Atmosphere":
        a = Atmosphere.earthlike()
        return Atmosphere(a.scale_height, a.thickness, tuple(b * 3.0 for b in a.rayleigh), a.mie * 3.0, a.mie_g)

    def hue(self) -> float:
        """Hue of the Rayleigh colour, degrees (src/sim.rs:264-266)."""
        return float(hsv(np.asarray(self.rayleigh, np.float32))[0])

    def tuned(self, density: float, hue: float | None = None) -> "Atmosphere":
        """The air scaled by ``density`` (Rayleigh and Mie), its scattering
        colour turned to ``hue`` at the same saturation, value and luminance
        (src/sim.rs:268-284)."""
        rayleigh = self.rayleigh
        if hue is not None:
            original = np.asarray(rayleigh, np.float32)
            _, saturation, value = hsv(original)
            turned = rgb(np.float32(hue), saturation, value)
            weights = np.array([0.2126, 0.7152, 0.0722], np.float32)
            gain = np.float32(original @ weights) / max(np.float32(turned @ weights), np.float32(1e-30))
            rayleigh = tuple(float(c * gain) for c in turned)
        return Atmosphere(self.scale_height, self.thickness, tuple(b * density for b in rayleigh),
                          self.mie * density, self.mie_g)


def hsv(c: np.ndarray) -> tuple:
    """RGB to (hue degrees, saturation, value) in f32 (src/sim.rs:1165-1180)."""
    c = np.asarray(c, np.float32)
    hi, lo = c.max(), c.min()
    delta = hi - lo
    if delta <= 0.0:
        hue = np.float32(0.0)
    elif hi == c[0]:
        hue = np.float32(60.0) * np.float32(np.mod((c[1] - c[2]) / delta, np.float32(6.0)))
    elif hi == c[1]:
        hue = np.float32(60.0) * ((c[2] - c[0]) / delta + np.float32(2.0))
    else:
        hue = np.float32(60.0) * ((c[0] - c[1]) / delta + np.float32(4.0))
    return hue, (delta / hi if hi > 0.0 else np.float32(0.0)), hi


def rgb(hue, saturation, value) -> np.ndarray:
    """(hue degrees, saturation, value) to RGB in f32 (src/sim.rs:1182-1198)."""
    f = np.float32
    c = f(value) * f(saturation)
    h = f(np.mod(f(hue), f(360.0))) / f(60.0)
    x = c * (f(1.0) - abs(np.fmod(h, f(2.0)) - f(1.0)))
    sector = int(h)
    r, g, b = [(c, x, 0), (x, c, 0), (0, c, x), (0, x, c), (x, 0, c)][sector] if sector < 5 else (c, 0, x)
    m = f(value) - c
    return np.array([r + m, g + m, b + m], np.float32)


@dataclass
class Scene:
    """``sim::Scene`` (src/sim.rs:312-334)."""

    name: str
    bodies: list
    host: int
    default_target: int
    targets: list
    atmosphere: Atmosphere | None
    ground: int
    default_fov_deg: float
    initial_direction_deg: tuple | None
    observer_lat_deg: float
    observer_lon_deg: float
    observer_height_m: float
    initial_time_days: float | None
    day_length_days: float | None
    indices: dict = field(default_factory=dict)

    def positions(self, t: float) -> np.ndarray:
        """World positions of every body at time ``t`` (days), AU (src/sim.rs:336-343)."""
        out = np.zeros((len(self.bodies), 3))
        for i, b in enumerate(self.bodies):
            if b.parent is not None:
                out[i] = out[b.parent] + b.relative_position(t)
        return out

    def star_indices(self) -> list:
        """src/sim.rs:426-433."""
        return [i for i, b in enumerate(self.bodies) if b.kind == KIND_STAR]

    def index(self, id_or_name: str) -> int:
        if id_or_name in self.indices:
            return self.indices[id_or_name]
        for i, b in enumerate(self.bodies):
            if b.name == id_or_name:
                return i
        raise KeyError(id_or_name)


# -- parsing (src/config.rs) ---------------------------------------------------
_SYSTEM_KEYS = {"version", "name", "time_days", "day_length_days", "camera", "targets", "bodies"}
_CAMERA_KEYS = {"body", "latitude_deg", "longitude_deg", "height_m", "look_at", "fov_deg", "atmosphere",
                "atmosphere_density", "ground", "direction"}
_BODY_KEYS = {"id", "name", "kind", "parent", "radius_km", "albedo", "luminosity", "emission", "orbit", "rotation",
              "rings", "surface"}
_ORBIT_KEYS = {"semi_major_axis_au", "semi_major_axis_km", "period_days", "eccentricity", "inclination_deg",
               "ascending_node_deg", "periapsis_deg", "mean_anomaly_deg", "epoch_days"}
_RING_KEYS = {"inner_radius", "outer_radius", "optical_depth", "albedo", "banded"}
_SURFACE_KEYS = {"style", "seed", "secondary_albedo", "belt_albedo", "bands", "turbulence", "storms", "craters",
                 "ice_caps", "ocean", "clouds"}


def _fields(d, allowed, where):
    _ensure(isinstance(d, dict), f"{where} must be a mapping")
    unknown = set(d) - allowed
    _ensure(not unknown, f"{where}: unknown field(s) {sorted(unknown)}")
    return d


def _finite(x) -> bool:
    return isinstance(x, (int, float)) and not isinstance(x, bool) and math.isfinite(x)


def _rgb(value, what):
    _ensure(isinstance(value, (list, tuple)) and len(value) == 3 and all(_finite(v) and 0.0 <= v <= 1.0 for v in value),
            f"{what} must contain three values in [0, 1]")
    return tuple(float(np.float32(v)) for v in value)


def _surface(d, albedo) -> Surface:
    """SurfaceDef::build (src/config.rs:191-266)."""
    _fields(d, _SURFACE_KEYS, "surface")
    _ensure(d.get("style") in STYLE_CODES, f"unknown surface style {d.get('style')!r}")
    style = d["style"]
    gas, terran = style == "gas_giant", style == "terran"
    _ensure(gas or all(d.get(k) is None for k in ("bands", "turbulence", "storms")),
            "surface bands, turbulence and storms apply only to gas_giant")
    _ensure(not gas or all(d.get(k) is None for k in ("craters", "ice_caps")),
            "surface craters and ice_caps do not apply to gas_giant")
    _ensure(terran or all(d.get(k) is None for k in ("ocean", "clouds")), "surface ocean and clouds apply only to terran")
    f32 = np.float32
    peak = max(f32(1.0e-6), *[f32(c) for c in albedo])

    def deeper(scale, power):
        return tuple(float(f32(scale) * f32(c) * f32(np.power(f32(c) / peak, f32(power)))) for c in albedo)

    secondary = d.get("secondary_albedo", d.get("belt_albedo"))
    if secondary is None:
        secondary = {
            "gas_giant": deeper(0.8, 1.6),
            "cratered": deeper(0.55, 0.6),
            "rocky": deeper(0.6, 1.0),
            "icy": (albedo[0] * 0.75, albedo[1] * 0.5, albedo[2] * 0.35),
            "volcanic": (albedo[0] * 0.85, albedo[1] * 0.45, albedo[2] * 0.2),
            "terran": (0.03, 0.09, 0.2),
        }[style]
    secondary = _rgb(secondary, "surface secondary_albedo")

    def unit(name, default):
        v = d.get(name, default)
        _ensure(_finite(v) and 0.0 <= v <= 1.0, f"surface {name} must be between 0 and 1")
        return float(v)

    bands = d.get("bands", 7.0)
    _ensure(_finite(bands) and 1.0 <= bands <= 30.0, "surface bands must be between 1 and 30")
    storms = d.get("storms", 1)
    _ensure(isinstance(storms, int) and 0 <= storms <= 3, "surface storms must be at most 3")
    craters, ice_caps = {"gas_giant": (0.0, 0.0), "cratered": (0.8, 0.0), "rocky": (0.5, 0.06), "icy": (0.15, 0.0),
                         "volcanic": (0.5, 0.0), "terran": (0.0, 0.1)}[style]
    seed = d.get("seed", 0)
    _ensure(isinstance(seed, int) and seed >= 0, "surface seed must be a nonnegative integer")
    return Surface(STYLE_CODES[style], seed, secondary, float(bands), unit("turbulence", 0.5), storms,
                   unit("craters", craters), unit("ice_caps", ice_caps), unit("ocean", 0.65), unit("clouds", 0.4))


def _body(d, indices, bodies) -> tuple:
    """BodyDef::build (src/config.rs:488-652); returns the body and its lock target id."""
    _fields(d, _BODY_KEYS, f"body '{d.get('id')}'")
    kind = d.get("kind")
    _ensure(kind in (KIND_ANCHOR, KIND_STAR, KIND_PLANET, KIND_MOON), f"unknown body kind {kind!r}")
    radius_km = d.get("radius_km")
    _ensure(_finite(radius_km) and (radius_km == 0.0 if kind == KIND_ANCHOR else radius_km > 0.0),
            "radius_km must be positive (zero for anchors)")
    albedo = _rgb(d.get("albedo", [0.5] * 3), "albedo")
    emission = _rgb(d.get("emission", [1.0] * 3), "emission")
    luminosity = d.get("luminosity", 0.0)
    _ensure(_finite(luminosity) and luminosity >= 0.0, "luminosity must be finite and nonnegative")
    _ensure(kind == KIND_STAR or luminosity == 0.0, "only stars may have luminosity")
    parent = indices[d["parent"]] if d.get("parent") is not None else None
    orbit = d.get("orbit")
    _ensure((parent is not None) == (orbit is not None), "parent and orbit must be provided together")
    if orbit is not None:
        _fields(orbit, _ORBIT_KEYS, "orbit")
        au, km = orbit.get("semi_major_axis_au"), orbit.get("semi_major_axis_km")
        _ensure((au is None) != (km is None), "orbit requires exactly one of semi_major_axis_au or semi_major_axis_km")
        a = au if au is not None else km / AU_KM
        _ensure(_finite(a) and a > 0.0, "semi-major axis must be positive and finite")
        period = orbit.get("period_days")
        _ensure(_finite(period) and period > 0.0, "orbital period_days must be positive and finite")
        e = orbit.get("eccentricity", 0.0)
        _ensure(_finite(e) and 0.0 <= e < 1.0, "eccentricity must be in [0, 1)")
        angles = [orbit.get(k, 0.0) for k in ("inclination_deg", "ascending_node_deg", "periapsis_deg",
                                                "mean_anomaly_deg", "epoch_days")]
        _ensure(all(_finite(x) for x in angles), "orbit angles and epoch must be finite")
        _ensure(a * (1.0 - e) > bodies[parent].radius + radius_km / AU_KM, "orbit intersects the parent at periapsis")
        inc, node, peri, m0, epoch = angles
        elements = Elements(a, e, math.radians(inc), math.radians(node), math.radians(peri), math.radians(m0), epoch)
        period_days = float(period)
    else:
        elements, period_days = Elements(0.0), 0.0
    rotation = d.get("rotation", {"mode": "free", "period_days": 1.0})
    _ensure(isinstance(rotation, dict) and rotation.get("mode") in ("free", "tidal_lock"), "unknown rotation mode")
    lock_id = None
    if rotation["mode"] == "free":
        _fields(rotation, {"mode", "period_days", "obliquity_deg", "phase_deg"}, "rotation")
        spin = rotation.get("period_days")
        obliquity, phase = rotation.get("obliquity_deg", 0.0), rotation.get("phase_deg", 0.0)
        _ensure(_finite(spin) and spin != 0.0 and _finite(obliquity) and _finite(phase),
                "rotation period must be nonzero; rotation values must be finite")
        obliquity, spin_period, spin_phase, locked = math.radians(obliquity), float(spin), math.radians(phase), False
    else:
        _fields(rotation, {"mode", "to"}, "rotation")
        lock_id = rotation.get("to")
        _ensure(parent is not None or lock_id is not None, "tidal_lock requires an orbit and parent, or a moon to lock to")
        obliquity, spin_period, spin_phase, locked = elements.inc, period_days, elements.m0, True
    rings = None
    if d.get("rings") is not None:
        r = _fields(d["rings"], _RING_KEYS, "rings")
        _ensure(kind in (KIND_STAR, KIND_PLANET, KIND_MOON), "rings require a star, planet or moon")
        inner, outer = r.get("inner_radius"), r.get("outer_radius")
        _ensure(_finite(inner) and _finite(outer) and inner > 1.0 and outer > inner,
                "ring radii must satisfy 1 < inner_radius < outer_radius")
        tau = r.get("optical_depth")
        _ensure(_finite(tau) and tau >= 0.0, "ring optical_depth must be finite and nonnegative")
        rings = Rings(float(inner), float(outer), float(np.float32(tau)), _rgb(r.get("albedo"), "ring albedo"),
                      bool(r.get("banded", False)))
    surface = None
    if d.get("surface") is not None:
        _ensure(kind in (KIND_PLANET, KIND_MOON), "surface requires a planet or moon")
        surface = _surface(d["surface"], albedo)
    body = Body(d.get("name") or d["id"], kind, parent, elements, period_days, radius_km / AU_KM, albedo, emission,
                float(np.float32(luminosity)), obliquity, spin_period, spin_phase, rings, surface, locked)
    return body, lock_id


def parse(text: str) -> Scene:
    """A scene from YAML system format v1 (src/config.rs:298-486)."""
    import yaml

    try:
        system = yaml.safe_load(text)
    except yaml.YAMLError as exc:
        raise SystemError(f"YAML schema error: {exc}") from exc
    _fields(system, _SYSTEM_KEYS, "system")
    _ensure(system.get("version") == 1, f"unsupported version {system.get('version')}; expected 1")
    _ensure(isinstance(system.get("name"), str) and system["name"].strip(), "system name cannot be empty")
    defs = system.get("bodies") or []
    _ensure(2 <= len(defs) <= 256, "a system needs 2..=256 bodies")
    time_days = system.get("time_days")
    _ensure(time_days is None or _finite(time_days), "time_days must be finite")
    ids = set()
    for b in defs:
        _ensure(isinstance(b, dict) and isinstance(b.get("id"), str) and b["id"].strip() and b["id"] not in ids,
                f"empty or duplicate body id '{b.get('id')}'")
        ids.add(b["id"])
    _ensure(sum(1 for b in defs if b.get("parent") is None) == 1,
            "exactly one root body is required (use an anchor for a binary system)")
    for b in defs:
        _ensure(b.get("parent") is None or b["parent"] in ids, f"body '{b['id']}': unknown parent '{b.get('parent')}'")
    # Stable topological order: a moon may come before its planet in the file.
    pending = list(defs)
    indices, bodies, locks = {}, [], []
    while pending:
        nxt = next((k for k, b in enumerate(pending) if b.get("parent") is None or b["parent"] in indices), None)
        _ensure(nxt is not None, "cyclic parent references: " + ", ".join(b["id"] for b in pending))
        d = pending.pop(nxt)
        body, lock_id = _body(d, indices, bodies)
        if lock_id is not None:
            locks.append((len(bodies), d["id"], lock_id))
        indices[d["id"]] = len(bodies)
        bodies.append(body)
    for index, id_, moon in locks:
        _ensure(moon in indices, f"body '{id_}': unknown tidal_lock moon '{moon}'")
        target = indices[moon]
        _ensure(bodies[target].parent == index, f"body '{id_}': tidal_lock to '{moon}' requires '{moon}' to orbit '{id_}'")
        bodies[index].lock_to = (bodies[target].elements, bodies[target].period_days)
    _ensure(any(b.kind == KIND_STAR and b.luminosity > 0.0 and any(x > 0.0 for x in b.emission) for b in bodies),
            "at least one luminous star is required")
    c = _fields(system.get("camera"), _CAMERA_KEYS, "camera")
    _ensure(c.get("body") in indices, "camera.body: unknown body")
    host = indices[c["body"]]
    _ensure(bodies[host].kind in (KIND_PLANET, KIND_MOON), "camera.body must be a planet or moon")
    _ensure(c.get("look_at") in indices, "camera.look_at: unknown body")
    default_target = indices[c["look_at"]]

    def check_target(i):
        _ensure(i != host and bodies[i].kind != KIND_ANCHOR,
                f"target '{bodies[i].name}' must be a visible body other than the camera host")

    check_target(default_target)
    day = system.get("day_length_days")
    _ensure(day is None or (_finite(day) and day > 0.0), "day_length_days must be positive and finite")
    targets = []
    for id_ in system.get("targets") or []:
        _ensure(id_ in indices, f"unknown body '{id_}'")
        check_target(indices[id_])
        targets.append(indices[id_])
    targets = targets or [default_target]
    lat, lon = c.get("latitude_deg"), c.get("longitude_deg")
    _ensure(_finite(lat) and -90.0 <= lat <= 90.0, "camera.latitude_deg must be in [-90, 90]")
    _ensure(_finite(lon) and -180.0 <= lon <= 180.0, "camera.longitude_deg must be in [-180, 180]")
    height = c.get("height_m", 2.0)
    _ensure(_finite(height) and height >= 0.1, "camera.height_m must be at least 0.1 metres")
    density = c.get("atmosphere_density")
    _ensure(density is None or (_finite(density) and 0.0 <= density <= 10.0),
            "camera.atmosphere_density must be between 0 and 10")
    fov = c.get("fov_deg")
    _ensure(_finite(fov) and MIN_FOV_DEG <= fov <= MAX_FOV_DEG, "camera.fov_deg must be in [0.001, 90]")
    direction = c.get("direction")
    if direction is not None:
        _fields(direction, {"azimuth_deg", "altitude_deg"}, "camera.direction")
        az, alt = direction.get("azimuth_deg"), direction.get("altitude_deg")
        _ensure(_finite(az) and 0.0 <= az < 360.0, "camera.direction.azimuth_deg must be in [0, 360)")
        _ensure(_finite(alt) and -90.0 <= alt <= 90.0, "camera.direction.altitude_deg must be in [-90, 90]")
        direction = (float(az), float(alt))
    atmosphere = None
    if c.get("atmosphere") is not None:
        _ensure(c["atmosphere"] in ("earthlike", "dense"), f"unknown atmosphere {c['atmosphere']!r}")
        base = Atmosphere.earthlike() if c["atmosphere"] == "earthlike" else Atmosphere.dense()
        atmosphere = base.tuned(1.0 if density is None else float(density))
    ground = c.get("ground", "forest")
    _ensure(ground in THEMES, f"unknown ground {ground!r}")
    return Scene(system["name"], bodies, host, default_target, targets, atmosphere, THEMES[ground], float(fov),
                 direction, float(lat), float(lon), float(height), time_days, day, indices)


def load(path) -> Scene:
    with open(path, encoding="utf-8") as f:
        return parse(f.read())


# -- the observer (src/camera.rs) ----------------------------------------------
@dataclass
class Observer:
    """The telescope on a body's surface (src/camera.rs:7-22): radians."""

    body: int
    lat: float
    lon: float
    az: float = 0.0
    alt: float = math.radians(45.0)
    fov_y: float = math.radians(10.0)
    height_m: float = 2.0

    @staticmethod
    def for_scene(scene: Scene, t: float, lift_m: float = 0.0) -> "Observer":
        """The startup view: the file's direction, or aimed at its look_at body,
        with the eye ``lift_m`` above the configured height (the terrain under
        the lookout; src/main.rs:222-234, 600-625)."""
        o = Observer(scene.host, math.radians(scene.observer_lat_deg), math.radians(scene.observer_lon_deg),
                     height_m=scene.observer_height_m + lift_m)
        if scene.initial_direction_deg is not None:
            o.az, o.alt = (math.radians(v) for v in scene.initial_direction_deg)
            o.fov_y = math.radians(scene.default_fov_deg)
        else:
            o.point_at_fov(scene, t, scene.default_target, math.radians(scene.default_fov_deg))
        return o

    def point_at_fov(self, scene: Scene, t: float, target: int, fov: float) -> None:
        """src/main.rs:392-400."""
        positions = scene.positions(t)
        vf = self.frame(scene, t, positions)
        d = positions[target] - vf.position
        d /= np.linalg.norm(d)
        self.alt = math.asin(float(np.clip(d @ vf.zenith, -1.0, 1.0)))
        self.az = math.atan2(float(d @ vf.east), float(d @ vf.north))
        self.fov_y = min(max(fov, math.radians(MIN_FOV_DEG)), math.radians(MAX_FOV_DEG))

    def frame(self, scene: Scene, t: float, positions: np.ndarray) -> "ViewFrame":
        """src/camera.rs:70-120 (without an inertial reference orientation)."""
        host = scene.bodies[self.body]
        q = host.spin(t)
        slat, clat = math.sin(self.lat), math.cos(self.lat)
        slon, clon = math.sin(self.lon), math.cos(self.lon)
        zenith = q @ np.array([clat * clon, clat * slon, slat])
        zenith /= np.linalg.norm(zenith)
        north = q @ np.array([-slat * clon, -slat * slon, clat])
        east = q @ np.array([-slon, clon, 0.0])
        position = positions[self.body] + zenith * (host.radius + self.height_m / METRES_PER_AU)
        salt, calt = math.sin(self.alt), math.cos(self.alt)
        saz, caz = math.sin(self.az), math.cos(self.az)
        forward = calt * (caz * north + saz * east) + salt * zenith
        forward /= np.linalg.norm(forward)
        right = caz * east - saz * north
        right /= np.linalg.norm(right)
        up = np.cross(right, forward)
        stars = scene.star_indices()
        sun = 0.0
        if stars:
            d = positions[stars[0]] - position
            sun = float(zenith @ (d / np.linalg.norm(d)))
        return ViewFrame(position, forward, right, up, zenith, north, east, sun)


@dataclass
class ViewFrame:
    position: np.ndarray
    forward: np.ndarray
    right: np.ndarray
    up: np.ndarray
    zenith: np.ndarray
    north: np.ndarray
    east: np.ndarray
    sun_altitude: float = 0.0  # sine of the first star's altitude

    def world_to_view(self, v: np.ndarray) -> np.ndarray:
        """Camera-relative world vector to telescope space, forward -Z (src/camera.rs:58-65)."""
        return np.array([v @ self.right, v @ self.up, -(v @ self.forward)])


# -- the frame (src/main.rs:822-1018) --------------------------------------------
def surface_height(terrain, radius_m: float):
    """Height of a heightfield ``(heights, cells)`` in the ``terrain`` module's
    layout at east/north metres on a host of ``radius_m``, scaled as
    ``terrain::scale`` scales the reference's (src/terrain.rs:79-81)."""
    heights, cells = terrain
    s = tr.scale(radius_m)

    def height(x: float, z: float) -> float:
        return float(tr.heights_at(heights, cells, np.array([x / s]), np.array([z / s]))[0]) * s
    return height


def build_frame(scene: Scene, t: float, width: int, height: int, observer: Observer | None = None,
                density: int = 100, ground: bool = True, air_density: float = 1.0, air_hue: float | None = None,
                terrain=None) -> Frame:
    """``build_frame_with_vegetation``: every visible body in telescope space,
    its centre split into f32 high and low parts after an f64 rotation,
    stars' radiance from their luminosity, rings, surface styles and spin
    axes, the host's air shell (removed on the airless crater ground, scaled
    and turned by ``air_density`` and ``air_hue``), and the ground scenery of
    the scene's theme seated on ``terrain`` (``(heights, cells)``, or a
    levelled sphere when None). Without an ``observer`` the scene's startup
    view is used, its eye lifted onto the terrain."""
    host_m = scene.bodies[scene.host].radius * METRES_PER_AU
    lay = surface_height(terrain, host_m) if terrain is not None else None
    if observer is None:
        observer = Observer.for_scene(scene, t, lay(0.0, 0.0) if lay and ground else 0.0)
    positions = scene.positions(t)
    vf = observer.frame(scene, t, positions)
    tan_half = math.tan(observer.fov_y * 0.5)
    records, atmosphere_host = [], 0
    f32 = np.float32
    for i, b in enumerate(scene.bodies):
        if b.kind == KIND_ANCHOR:
            continue
        if i == scene.host:
            atmosphere_host = len(records)
        view = vf.world_to_view(positions[i] - vf.position)
        center = view.astype(np.float32)
        low = (view - center.astype(np.float64)).astype(np.float32)
        star = b.kind == KIND_STAR
        radiance = f32(SUN_LIGHT * float(b.luminosity) / b.radius ** 2)
        r = np.zeros((), SPHERE)
        r["center"] = center
        r["radius"] = f32(b.radius)
        r["color"] = [f32(c) * radiance for c in b.emission] if star else b.albedo
        r["emissive"] = 1.0 if star else 0.0
        r["center_low"] = [*low, 0.0 if star else 1.0]
        spin = b.spin(t)
        if b.rings is not None:
            pole = vf.world_to_view(spin @ np.array([0.0, 0.0, 1.0]))
            r["ring_plane"] = [*pole, b.rings.inner_radius]
            r["ring_params"] = [b.rings.outer_radius, b.rings.optical_depth, 1.0 if b.rings.banded else 0.0, 0.0]
            r["ring_color"] = [*b.rings.albedo, 0.0]
        ax = vf.world_to_view(spin @ np.array([1.0, 0.0, 0.0]))
        ay = vf.world_to_view(spin @ np.array([0.0, 1.0, 0.0]))
        if b.surface is not None:
            s = b.surface
            gas = s.style == 1
            style, seed = float(s.style), s.seed
            r["surface_color"] = [*s.secondary_albedo, s.turbulence if gas else s.craters]
            r["surface_params"] = [s.bands, float(s.storms), 0.0, 0.0] if gas else [s.ice_caps, s.ocean, s.clouds, 0.0]
        else:
            style, seed = 0.0, i
        r["axis_x"] = [*ax, style]
        r["axis_y"] = [*ay, float(seed % 4096)]
        records.append(r)
    bodies = np.array(records, SPHERE)
    host = scene.bodies[scene.host]
    host_view = vf.world_to_view(positions[scene.host] - vf.position)
    g = np.zeros((), GLOBALS)
    g["cam_right"] = [*vf.right, 0.0]
    g["cam_up"] = [*vf.up, 0.0]
    g["cam_forward"] = [*vf.forward, tan_half]
    g["viewport"] = [width, height, 1.0, 0.0]
    atmosphere = scene.atmosphere
    if atmosphere is not None and ground and scene.ground == THEMES["craters"]:
        atmosphere = None  # the crater field is airless
    if atmosphere is not None and air_density > 0.005:
        a = atmosphere.tuned(air_density, air_hue)
        g["atmo_center"] = [*host_view, host.radius]
        g["atmo_rayleigh"] = [*a.rayleigh, a.mie]
        g["atmo_params"] = [a.mie_g, a.scale_height, a.thickness, atmosphere_host + 1]
    if ground:
        if scene.ground == gr.THEME_FOREST:
            prims = gr.landscape_with_density(host_m, density, lay)
        elif scene.ground in (gr.THEME_ALPINE, gr.THEME_DUNES):
            prims = np.zeros(0, PRIMITIVE)
        else:
            prims = gr.theme_rocks(host_m, scene.ground, lay)
    else:
        prims = np.zeros(0, PRIMITIVE)
    g["ground_east"] = [*vf.world_to_view(vf.east), observer.height_m]
    g["ground_up"] = [*vf.world_to_view(vf.zenith), host_m]
    g["ground_north"] = [*vf.world_to_view(vf.north), atmosphere_host + 1 if ground else 0.0]
    g["ground_counts"] = [len(prims), 1 if ground else 0, scene.ground, 0]
    return Frame(g, bodies, prims, float(t))


# -- the built-in binary system (src/sim.rs:452-783) ---------------------------
def binary_scene() -> Scene:
    """``sim::binary_scene``: two suns about a barycentre, the observatory on
    Calyx (35 N, 0 E) with Earth-like air, Halo the startup target. The
    reference builds it in code rather than from a file; these are its
    bodies (angles other than inclination and obliquity in radians)."""
    deg = math.pi / 180.0

    def body(name, kind, parent, el, period, radius_km, albedo, obliquity, spin):
        return Body(name, kind, parent, Elements(*el), period, radius_km / AU_KM, albedo, (0.0, 0.0, 0.0), 0.0,
                    obliquity, spin, 0.0)

    def star(name, a, m0, radius_au, emission, luminosity, spin):
        return Body(name, KIND_STAR, 0, Elements(a, 0.0, 0.0, 0.0, 0.0, m0), 120.0, radius_au, (0.0, 0.0, 0.0),
                    tuple(float(np.float32(c)) for c in emission), float(np.float32(luminosity)), 0.0, spin, 0.0)

    f = lambda *c: tuple(float(np.float32(x)) for x in c)  # noqa: E731  (f32 colours)
    bodies = [
        Body("Barycentre", KIND_ANCHOR, None, Elements(0.0), 0.0, 0.0, (0.0,) * 3, (0.0,) * 3, 0.0, 0.0, 1.0, 0.0),
        star("Aur", 0.30, 0.0, 0.030, (1.0, 0.96, 0.86), 2.0, 22.0),
        star("Igni", 0.50, math.pi, 0.018, (1.0, 0.42, 0.20), 0.7, 30.0),
        body("Nerid", KIND_PLANET, 0, (2.6, 0.03, 1.5 * deg, 0.4, 1.1, 2.0), 1530.0, 5_734.0, f(0.45, 0.55, 0.65),
             18.0 * deg, 0.85),
        body("Pip", KIND_MOON, 3, (2.2e-4, 0.01, 2.0 * deg, 0.0, 0.5, 1.0), 2.5, 1_200.0, f(0.60, 0.60, 0.62),
             2.0 * deg, 2.5),
        body("Calyx", KIND_PLANET, 0, (4.2, 0.02, 1.0 * deg, 2.0, 3.0, 4.5), 3000.0, 9_560.0, f(0.72, 0.60, 0.42),
             22.0 * deg, 0.60),
        body("Halo", KIND_MOON, 5, (1.6e-4, 0.002, 0.5 * deg, 0.0, 0.0, 0.0), 3.5, 1_500.0, f(0.55, 0.57, 0.60),
             0.5 * deg, 3.5),
        body("Umbra", KIND_MOON, 5, (3.2e-4, 0.01, 1.5 * deg, 1.0, 2.0, 3.0), 8.0, 1_000.0, f(0.35, 0.33, 0.32),
             1.5 * deg, 8.0),
        body("Vantus", KIND_PLANET, 0, (6.8, 0.05, 2.5 * deg, 1.5, 0.8, 1.2), 6200.0, 14_016.0, f(0.35, 0.42, 0.55),
             12.0 * deg, 0.40),
        body("Iri", KIND_MOON, 8, (1.8e-4, 0.005, 1.0 * deg, 0.0, 1.0, 0.5), 5.0, 800.0, f(0.50, 0.50, 0.55),
             1.0 * deg, 5.0),
        body("Kesh", KIND_MOON, 8, (3.0e-4, 0.01, 2.0 * deg, 0.5, 2.0, 2.5), 11.0, 600.0, f(0.45, 0.40, 0.35),
             2.0 * deg, 11.0),
        body("Coll", KIND_MOON, 8, (4.5e-4, 0.02, 3.0 * deg, 0.2, 4.0, 5.0), 18.0, 500.0, f(0.30, 0.32, 0.38),
             3.0 * deg, 18.0),
    ]
    return Scene("Binary", bodies, 5, 6, [6, 1, 2, 7, 3, 8, 4, 9], Atmosphere.earthlike(), THEMES["forest"], 20.0,
                 None, 35.0, 0.0, 2.0, None, None)
