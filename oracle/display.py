"""ORACLE — test infrastructure only: a numpy restatement of the reference's
display-only exposure and tone mapping, the checker for the HIP kernels
k_meter, k_resolve and k_display. Nothing in the product imports it.

- ``meter_size``: src/pathtracer.rs:40-42.
- ``histogram``: the measure pass, src/shaders/exposure.wgsl:51-76.
- ``resolve``: src/shaders/exposure.wgsl:78-109.
- ``display``: the fragment shader src/shaders/display.wgsl:36-97 at each pixel
  centre of an out_w x out_h target (uv = (x + 0.5) / out_w, (y + 0.5) / out_h,
  top row first), in float32.

Pinned by the reference's own tests restated in tests/test_display.py: the
meter rectangle (src/performance_tests.rs:193-209), the upscale orientation
and edges (src/pathtracer_tests.rs:318-435), extreme exposures
(:1268-1360) and single-pixel metering (src/ring_tests.rs:454-472).
"""
from __future__ import annotations

import numpy as np

BINS = 256
LOG_MIN, LOG_MAX = -24.0, 24.0
METER_FLOOR, MIN_EXPOSURE, MAX_EXPOSURE = 1.0e-3, 1.0e-8, 1.0e8
f32 = np.float32


def meter_size(w: int, h: int) -> tuple[int, int]:
    return w - 2 * (w // 4), h - 2 * (h // 4)


def histogram(accumulation: np.ndarray) -> np.ndarray:
    """256 counts over the central rectangle of an (H, W, 4) f32 image."""
    h, w = accumulation.shape[:2]
    ox, oy = w // 4, h // 4
    mw, mh = meter_size(w, h)
    hdr = accumulation[oy:oy + mh, ox:ox + mw].astype(f32)
    with np.errstate(invalid="ignore", over="ignore"):
        lum = hdr[..., 0] * f32(0.2126) + hdr[..., 1] * f32(0.7152) + hdr[..., 2] * f32(0.0722) - hdr[..., 3]
        lum = np.maximum(f32(0.0), lum)  # NaN stays NaN here, as max(0, NaN) may
        safe = np.where(lum > 0, lum, f32(2.0 ** LOG_MIN)).astype(f32)
        safe = np.clip(safe, f32(2.0 ** LOG_MIN), f32(2.0 ** LOG_MAX))
        t = (np.log2(safe).astype(f32) - f32(LOG_MIN)) / f32(LOG_MAX - LOG_MIN)
        b = np.clip(t * f32(BINS), 0, BINS - 1).astype(np.uint32)
    return np.bincount(b.ravel(), minlength=BINS).astype(np.uint32)


def resolve(hist: np.ndarray, value: float, key: float, percentile: float, enabled: bool) -> float:
    """The eased exposure after one metering."""
    total = int(hist.sum())
    if not enabled or total == 0:
        return value
    cutoff = max(1, int(np.ceil(f32(total) * f32(min(max(percentile, 0.0), 1.0)))))
    bin_ = int(np.argmax(np.cumsum(hist) >= cutoff))
    t = (bin_ + 0.5) / BINS
    measured = max(2.0 ** (LOG_MIN + t * (LOG_MAX - LOG_MIN)), METER_FLOOR)
    desired = min(max(key / measured, MIN_EXPOSURE), MAX_EXPOSURE)
    return min(max(value + (desired - value) * 0.15, MIN_EXPOSURE), MAX_EXPOSURE)


def aces(x):
    x = np.asarray(x, f32)
    return np.clip((x * (f32(2.51) * x + f32(0.03))) / (x * (f32(2.43) * x + f32(0.59)) + f32(0.14)), 0.0, 1.0)


def display(accumulation, backdrop, out_w, out_h, exposure, auto, auto_value, bias, milky_way):
    """Mapped linear display values (out_h, out_w, 3)."""
    acc = np.asarray(accumulation, f32)
    sky_img = np.asarray(backdrop, f32)[..., :3]
    h, w = acc.shape[:2]
    xs = ((np.arange(out_w, dtype=f32) + f32(0.5)) / f32(out_w) * f32(w) - f32(0.5)).astype(f32)
    ys = ((np.arange(out_h, dtype=f32) + f32(0.5)) / f32(out_h) * f32(h) - f32(0.5)).astype(f32)
    lx, ly = np.floor(xs).astype(int), np.floor(ys).astype(int)
    fx, fy = (xs - np.floor(xs)).astype(f32), (ys - np.floor(ys)).astype(f32)

    def at(img, y, x):
        return img[np.clip(y, 0, h - 1)[:, None], np.clip(x, 0, w - 1)[None, :]]

    def bilinear(img):
        a = at(img, ly, lx) + (at(img, ly, lx + 1) - at(img, ly, lx)) * fx[None, :, None]
        b = at(img, ly + 1, lx) + (at(img, ly + 1, lx + 1) - at(img, ly + 1, lx)) * fx[None, :, None]
        return a + (b - a) * fy[:, None, None]

    hdr = bilinear(acc)[..., :3]
    sky = bilinear(sky_img)
    base = auto_value if auto else exposure
    stops = np.clip(np.log2(np.clip(f32(base), 1e-30, 1e30)) + np.clip(bias, -100, 100), -100, 100)
    multiplier = f32(2.0 ** stops)
    foreground = np.maximum(hdr - sky, f32(0.0))
    gain = min(multiplier, f32(0.04) / max(f32(milky_way), f32(1e-30)))
    x = np.minimum(foreground, f32(1.0e6) / multiplier) * multiplier + sky * gain
    return aces(x)


def srgb8(v):
    v = np.asarray(v, np.float64)
    e = np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055)
    return np.clip(np.floor(e * 255.0 + 0.5), 0, 255).astype(np.uint8)
