"""Deterministic offscreen renders from a versioned JSON request.

Restates the reference's ``stargaze render --request FILE --output PNG``
(src/headless.rs:14-216): the request (version 1: a scene file relative to
the request, a quality mode, the display size, a sample count and a manual
exposure) with the same checks and messages, the quality presets of
``graphics::Settings`` (src/graphics.rs:27-91), the startup view of the
scene file (``State::with_scene_settings(scene, false)``), progressive passes
up to the sample count, an sRGB PNG of the displayed image, and the
``Capture`` JSON line it prints.

Usage: ``python -m stargaze_amd.headless --request FILE --output PNG``.
"""
from __future__ import annotations

import json
import sys
import time
from dataclasses import asdict, dataclass
from pathlib import Path

from . import render_benchmark as rb
from . import stars as st
from . import system as sy
from . import terrain as tr
from .pathtracer import Options, PathTracer


@dataclass(frozen=True)
class Settings:
    """``graphics::Settings`` (src/graphics.rs:27-37)."""

    max_fps: int
    resolution_percent: int
    max_bounces: int
    samples_per_frame: int
    still_samples: int
    hud_scale_percent: int
    vegetation_percent: int

    def render_size(self, size: tuple[int, int]) -> tuple[int, int]:
        """The traced size at this preset's resolution (src/graphics.rs:87-91)."""
        scale = min(max(self.resolution_percent, 25), 100)
        return tuple(max(-(-v * scale // 100), 1) for v in size)


# Eco, Balanced, High (src/graphics.rs:46-74).
PRESETS = {
    "eco": Settings(15, 50, 2, 1, 128, 100, 35),
    "balanced": Settings(30, 75, 4, 1, 256, 100, 70),
    "high": Settings(60, 100, 8, 1, 1024, 100, 100),
}
REQUEST_FIELDS = {"version", "scene", "mode", "width", "height", "samples", "exposure", "adapter"}


class RequestError(ValueError):
    """An invalid render request (the reference's ``anyhow`` errors)."""


@dataclass
class Request:
    """``Request`` (src/headless.rs:34-45)."""

    version: int
    scene: str
    mode: str
    width: int
    height: int
    samples: int
    exposure: float
    adapter: str = ""

    @staticmethod
    def parse(text: str) -> "Request":
        d = json.loads(text)
        if not isinstance(d, dict):
            raise RequestError("a render request is a JSON object")
        unknown = set(d) - REQUEST_FIELDS
        if unknown:
            raise RequestError(f"unknown field(s) {sorted(unknown)}")
        missing = REQUEST_FIELDS - {"adapter"} - set(d)
        if missing:
            raise RequestError(f"missing field(s) {sorted(missing)}")
        if d["mode"] not in PRESETS:
            raise RequestError(f"unknown mode {d['mode']!r}")
        return Request(**d)

    def validate(self) -> None:
        """src/headless.rs:93-110."""
        if self.version != 1:
            raise RequestError(f"unsupported render request version {self.version}")
        if not (16 <= self.width <= 3840 and 16 <= self.height <= 2160):
            raise RequestError("render dimensions must be 16–3840 wide and 16–2160 high")
        if not 1 <= self.samples <= 4096:
            raise RequestError("samples must be 1–4096")
        exposure = float(self.exposure)
        if not (exposure == exposure and abs(exposure) != float("inf") and exposure > 0.0):
            raise RequestError("exposure must be finite and positive")


def load_scene(request: Request, request_path: Path) -> sy.Scene:
    """The request's scene file, relative to the request (src/headless.rs:111-119)."""
    scene = sy.load(Path(request_path).parent / request.scene)
    if scene.initial_time_days is None:
        raise RequestError("render scenes require an explicit time_days")
    return scene


def render(request: Request, scene: sy.Scene, output: Path, terrain_for=tr.for_theme) -> dict:
    """Render ``request`` into the PNG ``output``; returns the Capture record
    (src/headless.rs:123-216)."""
    preset = PRESETS[request.mode]
    terrain = terrain_for(scene.ground)
    radius_m = scene.bodies[scene.host].radius * sy.METRES_PER_AU
    observer = sy.Observer.for_scene(scene, scene.initial_time_days,
                                     sy.surface_height(terrain, radius_m)(0.0, 0.0))
    frame = sy.build_frame(scene, scene.initial_time_days, request.width, request.height, observer=observer,
                           density=preset.vegetation_percent, terrain=terrain)
    trace_size = preset.render_size((request.width, request.height))
    catalogue, cells = st.ray_catalogue(st.generate(*rb.STAR_CATALOGUE))
    tracer = PathTracer(trace_size, catalogue=catalogue, sky_cells=cells,
                        options=Options(samples_per_frame=preset.samples_per_frame, max_bounces=preset.max_bounces,
                                        sample_limit=request.samples, exposure=float(request.exposure)))
    tracer.upload_terrain(tr.terrain_data(*terrain), terrain[1])
    tracer.synchronize()
    start = time.perf_counter()
    while tracer.samples() < request.samples:
        tracer.encode(frame)
        tracer.synchronize()
    render_ms = (time.perf_counter() - start) * 1e3
    rb.write_png(Path(output), tracer.display_srgb8(request.width, request.height), srgb=True)
    adapter = rb._adapter()
    capture = {"version": 1, "mode": request.mode, "width": request.width, "height": request.height,
               "trace_width": trace_size[0], "trace_height": trace_size[1], "samples": tracer.samples(),
               "exposure": float(request.exposure), "render_ms": render_ms, "preset": asdict(preset),
               "gpu": adapter["name"], "backend": adapter["backend"], "driver": adapter["driver"],
               "driver_info": adapter["driver_info"]}
    tracer.close()
    return capture


def run(argv) -> int:
    """``stargaze render --request FILE --output PNG`` (src/headless.rs:65-121)."""
    slots = {"--request": None, "--output": None}
    args = list(argv)
    while args:
        arg = args.pop(0)
        if arg in ("--help", "-h"):
            print("Usage: python -m stargaze_amd.headless --request FILE --output PNG")
            return 0
        if arg not in slots:
            raise RequestError(f"unknown render argument: {arg}")
        if slots[arg] is not None:
            raise RequestError(f"duplicate render argument: {arg}")
        if not args:
            raise RequestError("missing render argument value")
        slots[arg] = Path(args.pop(0))
    if slots["--request"] is None:
        raise RequestError("render requires --request FILE")
    if slots["--output"] is None:
        raise RequestError("render requires --output PNG")
    request = Request.parse(slots["--request"].read_text())
    request.validate()
    scene = load_scene(request, slots["--request"])
    print(json.dumps(render(request, scene, slots["--output"])))
    return 0


if __name__ == "__main__":
    sys.exit(run(sys.argv[1:]))
