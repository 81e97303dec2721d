"""A headless render request end to end on the HIP path: the PNG and the
Capture record the reference prints (src/headless.rs:123-216)."""
import json

import numpy as np
import pytest

from stargaze_amd import headless as hl
from stargaze_amd import system as sy

pytestmark = pytest.mark.gpu


def test_headless_request_renders_png_and_capture(tmp_path, capsys):
    (tmp_path / "vesper.yaml").write_text((sy.CONFIGS / "vesper.yaml").read_text())
    req = {"version": 1, "scene": "vesper.yaml", "mode": "eco", "width": 64, "height": 36, "samples": 4,
           "exposure": 2.0}
    (tmp_path / "request.json").write_text(json.dumps(req))
    print("headless: vesper, eco 64x36, 4 spp", flush=True)
    assert hl.run(["--request", str(tmp_path / "request.json"), "--output", str(tmp_path / "out.png")]) == 0
    capture = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert (capture["version"], capture["mode"], capture["width"], capture["height"]) == (1, "eco", 64, 36)
    assert (capture["trace_width"], capture["trace_height"], capture["samples"]) == (32, 18, 4)
    assert capture["exposure"] == 2.0 and capture["render_ms"] > 0.0
    assert capture["preset"]["max_bounces"] == 2 and capture["preset"]["vegetation_percent"] == 35
    assert set(capture) == {"version", "mode", "width", "height", "trace_width", "trace_height", "samples",
                            "exposure", "render_ms", "preset", "gpu", "backend", "driver", "driver_info"}
    png = (tmp_path / "out.png").read_bytes()
    assert png[:8] == b"\x89PNG\r\n\x1a\n" and b"sRGB" in png[:64]
