"""The display/exposure oracle against the reference's own pins (host only)."""
import math

import numpy as np

import display as od


def test_meter_rectangle_matches_every_pixel_of_the_central_region():
    """src/performance_tests.rs:193-209."""
    for w in range(1, 130):
        for h in range(1, 80):
            expected = (sum(1 for x in range(w) if w // 4 <= x < w - w // 4),
                        sum(1 for y in range(h) if h // 4 <= y < h - h // 4))
            assert od.meter_size(w, h) == expected
    assert od.meter_size(3840, 2160) == (1920, 1080)


def test_upscale_orientation_and_edges():
    """src/pathtracer_tests.rs:318-435: a 2x2 trace of red, green, blue and
    white shown at 4x4, 2x2 and 1x1."""
    acc = np.array([[[1, 0, 0, 0], [0, 1, 0, 0]], [[0, 0, 1, 0], [1, 1, 1, 0]]], np.float32)
    zero = np.zeros_like(acc)
    for w, h in ((4, 4), (2, 2), (1, 1)):
        out = od.display(acc, zero, w, h, 1.0, False, 1.0, 0.0, 0.0)
        if w == 1:
            np.testing.assert_allclose(out[0, 0], od.aces([0.5] * 3), atol=1e-5)
            continue
        np.testing.assert_allclose(out[0, 0], od.aces([1, 0, 0]), atol=1e-5)
        np.testing.assert_allclose(out[0, w - 1], od.aces([0, 1, 0]), atol=1e-5)
        np.testing.assert_allclose(out[h - 1, 0], od.aces([0, 0, 1]), atol=1e-5)
        np.testing.assert_allclose(out[h - 1, w - 1], od.aces([1, 1, 1]), atol=1e-5)
        if w == 4:
            np.testing.assert_allclose(out[1, 1], od.aces([0.625, 0.25, 0.25]), atol=1e-5)


def test_extreme_exposure():
    """src/pathtracer_tests.rs:1268-1360: an emitter of radiance 10."""
    acc = np.full((1, 1, 4), 10.0, np.float32)
    acc[..., 3] = 0.0
    for base, bias, expected in ((1.0, 80.0, 1.0), (1.0, -1000.0, 0.0), (1.0, 1000.0, 1.0),
                                 (float(np.finfo(np.float32).max), 0.0, 1.0)):
        out = od.display(acc, np.zeros_like(acc), 1, 1, base, False, 1.0, bias, 0.0)
        assert np.all(np.isfinite(out)) and np.all(np.abs(out - expected) < 1e-5), (base, bias, out)


def test_single_pixel_auto_exposure():
    """src/ring_tests.rs:454-472: a lone pixel of radiance 10 meters to
    key / 10 = 0.025 within 10% after 64 eased passes."""
    acc = np.full((1, 1, 4), 10.0, np.float32)
    acc[..., 3] = 0.0
    hist = od.histogram(acc)
    assert hist.sum() == 1
    value = 1.0
    for _ in range(64):
        value = od.resolve(hist, value, 0.25, 0.9, True)
    assert abs(value / 0.025 - 1) < 0.1, value


def test_histogram_special_values():
    """Zero, NaN and negative luminance fall into the darkest bin; the
    Milky Way luminance in w is subtracted; infinity clamps to the top bin."""
    acc = np.zeros((4, 4, 4), np.float32)
    acc[1, 1] = [np.nan, 0, 0, 0]
    acc[1, 2] = [-1, -1, -1, 0]
    acc[2, 1] = [np.inf, 0, 0, 0]
    acc[2, 2] = [1, 1, 1, 0]
    hist = od.histogram(acc)
    assert hist.sum() == 4
    assert hist[0] == 2 and hist[255] == 1
    one = int((math.log2(1.0) + 24) / 48 * 256)
    assert hist[one] == 1
