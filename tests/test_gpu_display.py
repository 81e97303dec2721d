"""Exposure metering and display on MI355X against the display oracle and
the reference's pins."""
import numpy as np
import pytest

import display as od
from stargaze_amd import scenes
from stargaze_amd.pathtracer import Options, PathTracer

pytestmark = pytest.mark.gpu


def emitter_frame(size):
    f = scenes.frame()
    f.globals["viewport"] = [size[0], size[1], 1.0, 0.0]
    f.bodies = scenes.bodies(scenes.body([0.0, 0.0, -3.0], 1.0, [10.0] * 3, True))
    return f


def special_image(w, h):
    """src/performance_tests.rs:219-233: zero, NaN, infinity, negative,
    Milky-Way-only, and a ladder of luminances over 32 stops."""
    out = np.zeros((h, w, 4), np.float32)
    for i in range(w * h):
        k = i % 17
        if k == 0:
            v = [0, 0, 0, 0]
        elif k == 1:
            v = [np.nan, 0, 0, 0]
        elif k == 2:
            v = [np.inf, 0, 0, 0]
        elif k == 3:
            v = [-1, -1, -1, 0]
        elif k == 4:
            v = [8, 8, 8, 8]
        else:
            x = 2.0 ** ((i % 80) * 0.4 - 16.1)
            v = [x, x * 0.7, x * 0.3, 0]
        out[i // w, i % w] = v
    return out


def metered(size, image):
    """Meter `image` through sg_encode once the sample limit is reached."""
    t = PathTracer(size, options=Options(samples_per_frame=1, max_bounces=1, sample_limit=1, milky_way=0.0,
                                         auto_exposure=True))
    f = emitter_frame(size)
    t.encode(f)  # the one sample the limit allows
    t.write_accumulation(image)
    t.encode(f)  # nothing to trace: metering only
    state, hist = t.read_exposure()
    return t, state, hist


@pytest.mark.parametrize("size", [(1, 1), (2, 3), (5, 7), (16, 16), (17, 19), (31, 33), (129, 71), (513, 287)])
def test_meter_histogram_matches_oracle(size):
    print(f"meter {size}", flush=True)
    image = special_image(*size)
    t, state, hist = metered(size, image)
    expected = od.histogram(image)
    mw, mh = od.meter_size(*size)
    assert int(hist.sum()) == mw * mh
    np.testing.assert_array_equal(hist, expected)
    t.close()


def test_single_pixel_auto_exposure():
    """src/ring_tests.rs:454-472."""
    t = PathTracer((1, 1), options=Options(samples_per_frame=1, max_bounces=1, milky_way=0.0, auto_exposure=True))
    f = emitter_frame((1, 1))
    for _ in range(64):
        t.encode(f)
    state, _ = t.read_exposure()
    assert abs(float(state["value"]) / 0.025 - 1) < 0.1, state
    t.close()


def test_auto_exposure_after_convergence_keeps_the_image():
    """src/ring_tests.rs:474-503: metering past the sample limit changes the
    exposure, never the image."""
    size = (8, 8)
    t = PathTracer(size, options=Options(samples_per_frame=1, max_bounces=1, sample_limit=1, milky_way=0.0))
    f = emitter_frame(size)
    t.encode(f)
    before = t.read_accumulation()
    t.set_exposure_controls(True, 0.0)
    for _ in range(64):
        t.encode(f)
    after = t.read_accumulation()
    state, _ = t.read_exposure()
    assert np.array_equal(before, after)
    assert t.samples() == 1
    assert abs(float(state["value"]) / 0.025 - 1) < 0.1, state
    t.close()


def test_display_upscale_orientation_and_edges():
    """src/pathtracer_tests.rs:318-435."""
    t = PathTracer((2, 2), options=Options(samples_per_frame=1, max_bounces=1, milky_way=0.0))
    t.encode(emitter_frame((2, 2)))
    colors = np.array([[[1, 0, 0, 0], [0, 1, 0, 0]], [[0, 0, 1, 0], [1, 1, 1, 0]]], np.float32)
    t.write_accumulation(colors)
    t.write_backdrop(np.zeros_like(colors))
    for w, h in ((4, 4), (2, 2), (1, 1)):
        out = t.display(w, h)
        assert np.all(out[..., 3] == 1.0)
        if w == 1:
            np.testing.assert_allclose(out[0, 0, :3], od.aces([0.5] * 3), atol=1e-5)
            continue
        np.testing.assert_allclose(out[0, 0, :3], od.aces([1, 0, 0]), atol=1e-5)
        np.testing.assert_allclose(out[0, w - 1, :3], od.aces([0, 1, 0]), atol=1e-5)
        np.testing.assert_allclose(out[h - 1, 0, :3], od.aces([0, 0, 1]), atol=1e-5)
        np.testing.assert_allclose(out[h - 1, w - 1, :3], od.aces([1, 1, 1]), atol=1e-5)
        if w == 4:
            np.testing.assert_allclose(out[1, 1, :3], od.aces([0.625, 0.25, 0.25]), atol=1e-5)
    t.close()


def test_display_extreme_exposure():
    """src/pathtracer_tests.rs:1268-1360; the accumulation survives."""
    for base, bias, expected in ((1.0, 80.0, 1.0), (1.0, -1000.0, 0.0), (1.0, 1000.0, 1.0),
                                 (float(np.finfo(np.float32).max), 0.0, 1.0)):
        t = PathTracer((1, 1), options=Options(samples_per_frame=1, max_bounces=1, milky_way=0.0, exposure=base))
        t.set_exposure_controls(False, bias)
        f = emitter_frame((1, 1))
        for _ in range(4):
            t.encode(f)
        out = t.display(1, 1)
        assert np.all(np.isfinite(out)) and np.all(np.abs(out[0, 0, :3] - expected) < 1e-5), (base, bias, out)
        assert t.samples() == 4
        t.close()


@pytest.mark.parametrize("out_size", [(96, 54), (40, 30), (200, 113)])
def test_display_matches_oracle_on_a_render(out_size):
    """A rendered ringed giant with a Milky-Way-like backdrop, shown at
    other sizes, manual and automatic exposure: float and 8-bit sRGB."""
    size = (96, 54)
    # The sample limit stops later passes from tracing: they only publish
    # the controls and meter, so the images below stay as written.
    t = PathTracer(size, options=Options(samples_per_frame=4, max_bounces=3, milky_way=2.0e-4, sample_limit=4))
    frame = scenes.ringed_giant(*size)
    t.encode(frame)
    rng = np.random.default_rng(5)
    acc = t.read_accumulation()
    sky = np.zeros_like(acc)
    sky[..., :3] = rng.uniform(0, 1e-4, acc[..., :3].shape)
    acc[..., :3] += sky[..., :3]
    t.write_accumulation(acc)
    t.write_backdrop(sky)
    for auto in (False, True):
        t.set_exposure_controls(auto, 0.5)
        t.encode(frame)  # publishes the controls (and meters when automatic)
        acc = t.read_accumulation()
        state, _ = t.read_exposure()
        out = t.display(*out_size)
        expected = od.display(acc, sky, out_size[0], out_size[1], 1.0, auto, float(state["value"]), 0.5, 2.0e-4)
        np.testing.assert_allclose(out[..., :3], expected, rtol=1e-5, atol=2e-6)
        srgb = t.display_srgb8(*out_size)
        assert np.abs(srgb[..., :3].astype(int) - od.srgb8(out[..., :3]).astype(int)).max() <= 1
        assert np.all(srgb[..., 3] == 255)
    t.close()
