"""The reference's render benchmark suite on this path tracer.

Restates ``render_benchmark.rs``: its twenty named cases (``fixture``,
src/render_benchmark.rs:339-605), each a sequence of frames advancing one
simulated second per step; the JSON run configuration and its checks
(``Config``, :174-250); the FNV-1a frame fingerprint (:742-756); and the run
itself (:771-996): per case, paused and advancing timings over repeats,
captures at a fixed exposure, and ``results.json`` in the reference's schema.

Three things differ, each recorded in the results:

* Frames with a ground stand on the reference's own terrain
  (``reference_terrain``), except the alpine ground, which needs the
  reference's elevation survey: without ``STARGAZE_TERRAIN_SURVEYS`` it
  stands on a seeded stand-in.
* The basalt ground (cases ``ground-basalt`` and ``level-1``) is not built:
  those cases are recorded with a ``skipped`` reason instead of timings.
* ``gpu_ms`` is the HIP-event time of the integrator kernels of one
  ``encode``; the reference's GPU timestamps also cover its display pass,
  which here runs only for captures.

Usage: ``python -m stargaze_amd.render_benchmark CONFIG.json`` (the
reference reads the same JSON from ``STARGAZE_BENCH_CONFIG``).
"""
from __future__ import annotations

import json
import math
import struct
import sys
import time
import zlib
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import stars as st
from . import system as sy
from . import terrain as tr
from .layouts import AU_KM, GLOBALS, PRIMITIVE, SPHERE
from .pathtracer import Frame, Options, PathTracer

F32 = np.float32

CASES = (
    "halo-rings", "earth-daylight", "earth-twilight", "median-dense", "vantus-eclipse", "vantus-airless",
    "dual-eclipse", "moonlit-air", "eclipse-umbra-edge", "twilight-afterglow", "eclipse-totality", "atacama-day",
    "atacama-night", "ground-craters", "ground-alpine", "ground-basalt", "ground-dunes", "level-1", "level-2",
    "level-3",
)  # src/render_benchmark.rs:14-35
VARIANTS = ("reference",)  # the shader experiments of render_experiments.rs are not built
NOT_BUILT = {"ground-basalt": "ground theme 4 (basalt) is not built",
             "level-1": "ground theme 4 (basalt) is not built"}

# Fixed manual exposure of the captures (src/render_benchmark.rs:899-915).
EXPOSURES = {"vantus-eclipse": 8.0, "vantus-airless": 8.0, "moonlit-air": 64.0, "level-1": 110.0, "level-2": 0.16,
             "level-3": 1.5, "ground-craters": 0.4, "ground-dunes": 0.3, "ground-alpine": 0.25, "ground-basalt": 0.6,
             "twilight-afterglow": 16.0, "eclipse-totality": 64.0}

# The Sun as seen from Earth in the solar-system scene (AU, radiance; :289-292).
SUN_DISTANCE = F32(0.9833)
SUN_RADIUS = F32(4.6547456e-3)
SUN_RADIANCE = np.array([115384.68, 109615.445, 101538.516], np.float32)
MOTION_NOISE_SEED = 0xE42DB137
STAR_CATALOGUE = (3500, 2500, 0x5EED1234)


@dataclass
class Config:
    """The run configuration (src/render_benchmark.rs:174-250)."""

    output: str
    width: int
    height: int
    scale: int
    bounces: int
    vegetation: int
    warmup: int
    frames: int
    repeats: int
    image_samples: int
    adapter: str = ""
    cases: list = field(default_factory=lambda: list(CASES))
    capture_motion: bool = False
    motion_reference_samples: int = 0
    variant: str = "reference"

    @staticmethod
    def from_json(text: str) -> "Config":
        d = json.loads(text)
        known = set(Config.__dataclass_fields__)
        unknown = set(d) - known
        if unknown:
            raise ValueError(f"unknown config field(s) {sorted(unknown)}")
        return Config(**d)

    def validate(self) -> None:
        def ensure(ok, message):
            if not ok:
                raise ValueError(message)

        ensure(self.variant in VARIANTS, "unknown shader variant")
        ensure(16 <= self.width <= 3840 and 16 <= self.height <= 2160, "invalid dimensions")
        ensure(25 <= self.scale <= 100 and 1 <= self.bounces <= 16, "invalid quality")
        ensure(self.vegetation <= 100 and 1 <= self.warmup <= 100, "invalid vegetation/warmup")
        ensure(2 <= self.frames <= 1000 and 1 <= self.repeats <= 20, "invalid frames/repeats")
        ensure(2 <= self.image_samples <= 4096, "invalid image samples")
        ensure(self.motion_reference_samples == 0 or 2 <= self.motion_reference_samples <= 4096,
               "invalid motion reference samples")
        ensure(len(self.cases) > 0 and all(c in CASES for c in self.cases), "unknown or empty case selection")
        ensure(len(set(self.cases)) == len(self.cases), "duplicate case selection")

    def trace_size(self) -> tuple[int, int]:
        return -(-self.width * self.scale // 100), -(-self.height * self.scale // 100)


# -- f32 vector helpers (glam Vec3 arithmetic) --------------------------------
def _v(*c) -> np.ndarray:
    return np.array(c, np.float32)


def _rad(deg: float) -> np.float32:
    return F32(deg) * (F32(math.pi) / F32(180.0))


def sphere(center, radius, color, emissive) -> np.ndarray:
    """A bare body record (src/render_benchmark.rs:272-287)."""
    r = np.zeros((), SPHERE)
    r["center"] = np.asarray(center, np.float32)
    r["radius"] = F32(radius)
    r["color"] = np.asarray(color, np.float32)
    r["emissive"] = F32(emissive)
    return r


def _camera(cfg: Config, tan_half: np.float32) -> np.ndarray:
    g = np.zeros((), GLOBALS)
    g["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    g["cam_up"] = [0.0, 1.0, 0.0, 0.0]
    g["cam_forward"] = [0.0, 0.0, -1.0, tan_half]
    g["viewport"] = [cfg.width, cfg.height, 1.0, 0.0]
    return g


def _earthlike_air(g: np.ndarray, center, radius) -> None:
    air = sy.Atmosphere.earthlike()
    g["atmo_center"] = [*center, radius]
    g["atmo_rayleigh"] = [*air.rayleigh, air.mie]
    g["atmo_params"] = [air.mie_g, air.scale_height, air.thickness, 1.0]


def synthetic_earth(cfg: Config, altitude_deg: float, half_fov_deg: float):
    """Camera 2 m above an Earth-size host with Earth-like air, looking
    ``altitude_deg`` above the horizon (src/render_benchmark.rs:294-337).
    Returns the globals, the host centre and ``local(azimuth, elevation)``,
    a direction in degrees right of the view azimuth and above the horizon."""
    km = F32(1.0 / AU_KM)
    a = _rad(altitude_deg)
    up = _v(0.0, np.cos(a), -np.sin(a))
    north = _v(0.0, np.sin(a), np.cos(a))
    center = -up * (F32(6371.002) * km)
    g = _camera(cfg, np.tan(_rad(half_fov_deg)))
    _earthlike_air(g, center, F32(6371.0) * km)
    g["ground_east"] = [1.0, 0.0, 0.0, 2.0]
    g["ground_up"] = [*up, F32(6_371_000.0)]
    g["ground_north"] = [*north, 1.0]
    ahead = -north

    def local(azimuth: float, elevation: float) -> np.ndarray:
        az, el = _rad(azimuth), _rad(elevation)
        return (ahead * np.cos(az) + _v(1.0, 0.0, 0.0) * np.sin(az)) * np.cos(el) + up * np.sin(el)

    return g, center, local


def _frame(g, bodies, dt) -> Frame:
    return Frame(g, np.array(bodies, SPHERE), np.zeros(0, PRIMITIVE), float(dt))


def _synthetic(name: str, step: int, cfg: Config, dt: float) -> Frame | None:
    """The four hand-built views (src/render_benchmark.rs:342-464)."""
    km = F32(1.0 / AU_KM)
    s = F32(step)
    if name == "dual-eclipse":
        g = _camera(cfg, np.tan(_rad(30.0)))
        planet = sphere([0.0, 0.0, -3.0], 1.0, [0.55, 0.65, 0.8], 0.0)
        planet["ring_plane"] = [0.0, 0.8, 0.6, 1.2]
        planet["ring_params"] = [1.9, 0.7, 1.0, 0.0]
        planet["ring_color"] = [0.8, 0.65, 0.4, 0.0]
        return _frame(g, [planet, sphere([-3.0, 2.0, -1.0], 0.55, [25.0, 20.0, 12.0], 1.0),
                          sphere([3.0, 2.0, -1.0], 0.45, [12.0, 18.0, 25.0], 1.0),
                          sphere([F32(-1.5) + s * F32(0.003), 1.0, -1.5], 0.3, [0.3] * 3, 0.0),
                          sphere([F32(1.5) - s * F32(0.003), 1.0, -1.5], 0.25, [0.3] * 3, 0.0)], dt)
    if name == "eclipse-umbra-edge":
        g, center, _ = synthetic_earth(cfg, 12.0, 30.0)
        return _frame(g, [sphere(center, F32(6371.0) * km, [0.0] * 3, 0.0),
                          sphere([0.0, 0.0, -1.0], 1.0e-5, [2.0e10] * 3, 1.0),
                          sphere([(F32(998.0) - s * F32(0.1)) * km, 0.0, F32(-100_000.0) * km], F32(1000.0) * km,
                                 [0.0] * 3, 0.0)], dt)
    if name == "eclipse-totality":
        g, center, local = synthetic_earth(cfg, 5.0, 40.0)
        sun = local(0.0, 60.0)
        drift = local(90.0, 0.0) * (s * F32(0.5) * km)
        return _frame(g, [sphere(center, F32(6371.0) * km, [0.2] * 3, 0.0),
                          sphere(sun * SUN_DISTANCE, SUN_RADIUS, SUN_RADIANCE, 1.0),
                          sphere(sun * (F32(100_000.0) * km) + drift, F32(513.0) * km, [0.0] * 3, 0.0)], dt)
    if name == "moonlit-air":
        g = _camera(cfg, np.tan(_rad(27.5)))
        radius = F32(6371.0 / AU_KM)
        center = _v(0.0, 0.0, radius + F32(0.002 / AU_KM))
        _earthlike_air(g, center, radius)
        return _frame(g, [sphere(center, radius, [0.0] * 3, 0.0),
                          sphere([F32(0.0004) + s * F32(1e-7), 0.0, -0.002], 0.00015, [0.8] * 3, 0.0),
                          sphere([0.3, 0.0, 1.0], 0.005, [100_000.0] * 3, 1.0)], dt)
    return None


def _state(scene: sy.Scene) -> sy.Observer:
    """The benchmark's observer: the scene's site and configured height, the
    default mount (src/render_benchmark.rs:252-270, src/camera.rs:24-36)."""
    return sy.Observer(scene.host, math.radians(scene.observer_lat_deg), math.radians(scene.observer_lon_deg),
                       height_m=scene.observer_height_m)


def _point_initial_view(scene: sy.Scene, o: sy.Observer, t: float) -> None:
    """src/main.rs:609-621."""
    if scene.initial_direction_deg is not None:
        o.az, o.alt = (math.radians(v) for v in scene.initial_direction_deg)
        o.fov_y = math.radians(scene.default_fov_deg)
    else:
        o.point_at_fov(scene, t, scene.default_target, math.radians(scene.default_fov_deg))


def _lookout(scene: sy.Scene, terrain) -> float:
    """``terrain::lookout_height`` on the caller's heightfield (metres)."""
    if terrain is None:
        return 0.0
    return sy.surface_height(terrain, scene.bodies[scene.host].radius * sy.METRES_PER_AU)(0.0, 0.0)


def scene_for(name: str) -> sy.Scene | None:
    """The system a case observes, before its view is set."""
    if name == "halo-rings":
        return sy.load(_config("halo.yaml"))
    if name in ("earth-daylight", "earth-twilight", "twilight-afterglow"):
        return sy.load(_config("solar-system.yaml"))
    if name == "median-dense":
        return sy.load(_config("median-resonance.yaml"))
    if name in ("vantus-eclipse", "vantus-airless"):
        return sy.binary_scene()
    if name in ("atacama-day", "atacama-night") or name.startswith("ground-"):
        return sy.load(_config("atacama.yaml"))
    if name.startswith("level-"):
        return sy.load(_config({"level-1": "1-halo-volcano.yaml", "level-2": "2-erg-dunes.yaml",
                                "level-3": "3-umbra-valley.yaml"}[name]))
    return None


def _config(name: str) -> Path:
    return sy.CONFIGS / name


def ground_theme(name: str) -> int | None:
    """The ground theme a case renders, or None for the hand-built views."""
    if name in ("dual-eclipse", "eclipse-umbra-edge", "eclipse-totality", "moonlit-air"):
        return None
    if name.startswith("ground-"):
        return sy.THEMES[name[len("ground-"):]]
    return scene_for(name).ground


def fixture(name: str, step: int, cfg: Config, terrain=None) -> Frame:
    """The frame of case ``name`` at ``step`` (one simulated second per step;
    src/render_benchmark.rs:339-605). ``terrain`` is ``(heights, cells)`` for
    the case's ground, or None for a levelled sphere."""
    dt = step / 86400.0
    frame = _synthetic(name, step, cfg, dt)
    if frame is not None:
        return frame
    scene = scene_for(name)
    times = {"halo-rings": dt, "earth-daylight": 0.599 + dt, "earth-twilight": 1.04 + dt,
             "twilight-afterglow": 1.0525 + dt, "median-dense": 0.25 + dt, "vantus-eclipse": 444.478 + dt,
             "vantus-airless": 444.478 + dt, "atacama-day": 0.14 + dt, "atacama-night": dt}
    if name.startswith("level-"):
        # The app's own start: configured time, initial view, lookout height.
        t = scene.initial_time_days
        o = sy.Observer.for_scene(scene, t, _lookout(scene, terrain))
        t += dt
    elif name.startswith("ground-"):
        t = 0.14 + dt
        o = _state(scene)
        scene.ground = sy.THEMES[name[len("ground-"):]]
        o.height_m = scene.observer_height_m + _lookout(scene, terrain)
    else:
        t = times[name]
        o = _state(scene)
    if name in ("halo-rings", "atacama-day", "atacama-night"):
        _point_initial_view(scene, o, t)
    elif name.startswith("ground-"):
        _point_initial_view(scene, o, t)
        o.az = {"ground-alpine": math.radians(88.0), "ground-basalt": math.radians(10.0),
                "ground-dunes": math.radians(50.0)}.get(name, o.az)
        o.alt = math.radians({"ground-dunes": -14.0, "ground-craters": -12.0, "ground-basalt": -9.0,
                              "ground-alpine": -5.0}.get(name, 2.0))
        o.fov_y = math.radians(60.0)
    elif name in ("earth-daylight", "earth-twilight"):
        o.az, o.alt, o.fov_y = 0.0, math.radians(3.0), math.radians(55.0)
    elif name == "twilight-afterglow":
        sun = next(i for i, b in enumerate(scene.bodies) if b.luminosity > 0.0)
        o.point_at_fov(scene, t, sun, math.radians(70.0))
        o.az -= math.radians(25.0)
        o.alt = math.radians(12.0)
    elif name == "median-dense":
        o.az, o.alt, o.fov_y = 0.0, math.radians(20.0), math.radians(65.0)
    elif name in ("vantus-eclipse", "vantus-airless"):
        positions = scene.positions(t)
        view = o.frame(scene, t, positions)
        distance = float(np.linalg.norm(positions[8] - view.position))
        o.point_at_fov(scene, t, 8, math.asin(scene.bodies[8].radius / distance) * 2.5)
        if name == "vantus-airless":
            scene.atmosphere = None
    return sy.build_frame(scene, t, cfg.width, cfg.height, observer=o, density=cfg.vegetation, terrain=terrain)


def fingerprint(frame: Frame) -> str:
    """FNV-1a (64-bit) over the frame's exact GPU inputs
    (src/render_benchmark.rs:742-756)."""
    h = 0xCBF29CE484222325
    data = (np.asarray(frame.globals, GLOBALS).tobytes() + np.asarray(frame.bodies, SPHERE).tobytes()
            + np.asarray(frame.ground, PRIMITIVE).tobytes() + struct.pack("<d", float(frame.scene_time)))
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


stand_in_terrain = tr.stand_in
reference_terrain = tr.for_theme


# -- captures -------------------------------------------------------------------
def write_png(path: Path, rgba: np.ndarray, srgb: bool = False) -> None:
    """An 8-bit RGBA PNG (no filtering, one zlib stream), with an sRGB chunk
    (perceptual intent) when ``srgb``."""
    h, w = rgba.shape[:2]
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(kind: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
                + (chunk(b"sRGB", b"\x00") if srgb else b"")
                + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def _luma(rgba: np.ndarray) -> np.ndarray:
    """Rec.709 weights on the displayed sRGB channels, rounded (:718-731)."""
    p = rgba.astype(np.uint32)
    return ((2126 * p[..., 0] + 7152 * p[..., 1] + 722 * p[..., 2] + 5000) // 10000).astype(np.uint8)


def _capture(tracer: PathTracer, cfg: Config, name: str, exposure: float, pixels=None) -> dict:
    """The displayed image, its bytes and the linear accumulation (:673-740)."""
    out = Path(cfg.output)
    if pixels is None:
        pixels = tracer.display_srgb8(cfg.width, cfg.height)
    linear = tracer.read_accumulation()
    if not (np.all(np.isfinite(linear)) and np.all(linear >= 0.0)):
        raise RuntimeError(f"nonfinite/negative HDR in {name}")
    write_png(out / f"{name}.png", pixels)
    (out / f"{name}.rgba").write_bytes(pixels.tobytes())
    (out / f"{name}.f32").write_bytes(linear.astype("<f4").tobytes())
    return {"name": name, "spp": tracer.samples(), "exposure": exposure}


# -- the run ---------------------------------------------------------------------
def _adapter() -> dict:
    """The device the library runs on, from the HIP runtime."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    name = ctypes.create_string_buffer(256)
    runtime, driver = ctypes.c_int(), ctypes.c_int()
    hip.hipDeviceGetName(name, 256, 0)
    hip.hipRuntimeGetVersion(ctypes.byref(runtime))
    hip.hipDriverGetVersion(ctypes.byref(driver))
    return {"name": name.value.decode(errors="replace"), "vendor": 0x1002, "device": 0,
            "device_type": "DiscreteGpu", "backend": "Hip", "driver": f"HIP driver {driver.value}",
            "driver_info": f"HIP runtime {runtime.value}"}


class _Tracers:
    """One tracer per option set of a case, with the catalogue and terrain."""

    def __init__(self, cfg: Config, catalogue, terrain):
        self.cfg, self.catalogue, self.terrain = cfg, catalogue, terrain

    def make(self, options: Options) -> PathTracer:
        cat, cells = self.catalogue
        t = PathTracer(self.cfg.trace_size(), catalogue=cat, sky_cells=cells, options=options)
        if self.terrain is not None:
            t.upload_terrain(tr.terrain_data(*self.terrain), self.terrain[1])
        return t


def _render(tracer: PathTracer, frame: Frame) -> dict:
    """One timed ``encode`` (:607-665)."""
    ms0, _ = tracer.kernel_time()
    start = time.perf_counter()
    tracer.encode(frame)
    tracer.synchronize()
    wall = (time.perf_counter() - start) * 1e3
    ms1, _ = tracer.kernel_time()
    return {"gpu_ms": ms1 - ms0, "wall_ms": wall, "spp": tracer.samples()}


def median(values) -> float:
    """The suite's median: the mean of the two middle values (:886-888)."""
    v = sorted(values)
    return (v[(len(v) - 1) // 2] + v[len(v) // 2]) * 0.5


def time_cases(cfg: Config, terrain_for=reference_terrain, log=print) -> list:
    """The timed half of the run (:840-893): per case, ``warmup`` untimed
    submissions, then ``repeats`` runs of ``frames`` timed ones, paused (the
    first frame again and again) and advancing (the next frame each time).
    Each case keeps its frames and tracer factory under private keys for the
    captures."""
    cfg.validate()
    catalogue = st.ray_catalogue(st.generate(*STAR_CATALOGUE))
    options = Options(samples_per_frame=1, max_bounces=cfg.bounces, auto_exposure=True)
    cases = []
    for name in cfg.cases:
        if name in NOT_BUILT:
            log(f"{name:20} skipped: {NOT_BUILT[name]}")
            cases.append({"name": name, "skipped": NOT_BUILT[name]})
            continue
        theme = ground_theme(name)
        terrain = terrain_for(theme) if theme is not None else None
        tracers = _Tracers(cfg, catalogue, terrain)
        frames = [fixture(name, i, cfg, terrain) for i in range(max(cfg.frames, cfg.warmup))]
        modes = []
        for moving in (False, True):
            tracer = tracers.make(options)
            tracer.set_profiling(True)
            for i in range(cfg.warmup):
                tracer.encode(frames[i if moving else 0])
            repeats = []
            for _ in range(cfg.repeats):
                tracer.restart()
                measurements = []
                for i in range(cfg.frames):
                    timing = _render(tracer, frames[i if moving else 0])
                    if timing["spp"] != (1 if moving else i + 1):
                        raise RuntimeError("unexpected history/sample behavior")
                    measurements.append(timing)
                repeats.append(measurements)
            tracer.close()
            mode = "advancing" if moving else "paused"
            log(f"{name:20} {mode:9}: {median(m['gpu_ms'] for r in repeats for m in r):.3f} ms GPU median, "
                f"{median(m['wall_ms'] for r in repeats for m in r):.3f} ms wall median")
            modes.append({"mode": mode, "repeats": repeats})
        cases.append({"name": name, "initial_time_days": frames[0].scene_time,
                      "frame_fingerprints": [fingerprint(f) for f in frames[:cfg.frames]], "modes": modes,
                      "_frames": frames, "_tracers": tracers})
    return cases


def run(cfg: Config, terrain_for=reference_terrain, log=print) -> dict:
    """Time and capture every selected case; write ``results.json`` into
    ``cfg.output`` and return it."""
    cfg.validate()
    Path(cfg.output).mkdir(parents=True, exist_ok=True)
    cases = time_cases(cfg, terrain_for, log)
    # Captures after all timings (:894-985).
    for case in cases:
        if "skipped" in case:
            continue
        name, frames, tracers = case["name"], case.pop("_frames"), case.pop("_tracers")
        exposure = EXPOSURES.get(name, 1.0)
        image_options = Options(samples_per_frame=1, max_bounces=cfg.bounces, exposure=exposure)
        tracer = tracers.make(image_options)
        images = []
        for i in range(1, cfg.image_samples + 1):
            tracer.encode(frames[0])
            if i in (1, cfg.image_samples):
                images.append(_capture(tracer, cfg, f"{name}-spp{i}", exposure))
        tracer.restart()
        capture_motion = cfg.capture_motion or cfg.motion_reference_samples != 0
        for i, frame in enumerate(frames[:cfg.frames]):
            tracer.encode(frame)
            checkpoint = i in (cfg.frames // 2, cfg.frames - 1)
            if not checkpoint and not capture_motion:
                continue
            pixels = tracer.display_srgb8(cfg.width, cfg.height)
            if capture_motion:
                (Path(cfg.output) / f"{name}-moving-{i}.luma").write_bytes(_luma(pixels).tobytes())
            if checkpoint:
                images.append(_capture(tracer, cfg, f"{name}-moving-{i}", exposure, pixels))
        tracer.close()
        reference = cfg.motion_reference_samples
        if reference:
            ref = tracers.make(Options(noise_seed=MOTION_NOISE_SEED, samples_per_frame=64,
                                       max_bounces=cfg.bounces, sample_limit=reference, exposure=exposure))
            for i, frame in enumerate(frames[:cfg.frames]):
                ref.restart()
                for _ in range(-(-reference // 64)):
                    ref.encode(frame)
                if ref.samples() != reference:
                    raise RuntimeError("unexpected motion reference spp")
                pixels = ref.display_srgb8(cfg.width, cfg.height)
                (Path(cfg.output) / f"{name}-motion-reference-{i}.luma").write_bytes(_luma(pixels).tobytes())
            ref.close()
        case["images"] = images
    result = {"suite_version": 1, "variant": cfg.variant, "adapter": _adapter(), "timestamp_period_ns": 1.0,
              "trace_size": list(cfg.trace_size()), "cases": cases}
    (Path(cfg.output) / "results.json").write_text(json.dumps(result, indent=2))
    return result


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m stargaze_amd.render_benchmark CONFIG.json", file=sys.stderr)
        return 2
    cfg = Config.from_json(Path(argv[0]).read_text())
    run(cfg)
    return 0


if __name__ == "__main__":
    sys.exit(main())
