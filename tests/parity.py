"""Image-parity statistics shared by the GPU tests and the bench.

The HIP path and the oracle consume the same counter-based RNG streams, so
most pixels agree to float rounding; where a float comparison flips (a
Russian-roulette or null-crossing draw near its threshold) the two follow
different but equally distributed paths. Parity is therefore checked two
ways: per 16x16 tile, the means must agree within 3 combined standard errors
(from per-pixel second moments), and a large fraction of pixels must agree
to 1e-3 relative.
"""
from __future__ import annotations

import numpy as np


def tile_z(mean_a, m2_a, mean_b, m2_b, tile=16):
    """|mu_a - mu_b| / sqrt(se_a^2 + se_b^2) per tile and channel.

    ``mean_*``: (H, W, 4) running means; ``m2_*``: (H, W, 4) running means of
    the squared rgb with the sample count in w.
    """
    out = []
    h, w = mean_a.shape[:2]
    for ys in range(0, h, tile):
        for xs in range(0, w, tile):
            sl = (slice(ys, ys + tile), slice(xs, xs + tile))
            za = []
            for mean, m2 in ((mean_a, m2_a), (mean_b, m2_b)):
                mu = mean[sl][..., :3].astype(np.float64)
                n = m2[sl][..., 3:4].astype(np.float64)
                var = np.maximum(m2[sl][..., :3].astype(np.float64) - mu * mu, 0.0) / np.maximum(n, 1.0)
                p = mu.shape[0] * mu.shape[1]
                za.append((mu.mean(axis=(0, 1)), var.sum(axis=(0, 1)) / (p * p)))
            (ma, va), (mb, vb) = za
            diff = np.abs(ma - mb)
            se = np.sqrt(va + vb)
            scale = np.maximum(np.abs(ma), np.abs(mb))
            z = np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(diff <= 1e-5 * scale + 1e-12, 0.0, np.inf))
            out.append(z)
    return np.array(out)


def pixel_agreement(a, b, rel=1e-3, abs_=1e-7):
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    ok = np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + abs_
    return float(ok.all(axis=-1).mean())


def rmse(a, b):
    """Linear-RGB RMSE, W excluded (scripts/benchmark.py:205-233)."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    return float(np.sqrt(np.mean(d * d)))


def assert_tiles(z, max_fraction_over_3=0.01, max_z=6.0):
    over = float((z > 3.0).mean())
    worst = float(np.max(z))
    assert over <= max_fraction_over_3 and worst <= max_z, f"tiles over 3 sigma: {over:.4f}, worst z {worst:.2f}"
    return over, worst


def assert_same_draws(actual, expected, exact=True):
    """Two kernels that run the same step functions on the same draws (fused
    vs wavefront, tail vs compaction, batched vs per view) give the same
    image. The library is built without multiply-add contraction, so the
    steps round alike in every kernel and the images are bit-identical.
    With ``exact=False`` (a ground scene with air, where the wavefront
    integrator adds the camera-air contributions of deferred shadow rays in
    a sum of their own) they agree to 1e-5 relative, or 1e-7 of the image's
    peak for faint pixels."""
    actual, expected = np.asarray(actual), np.asarray(expected)
    if exact:
        differ = actual != expected
        assert not differ.any(), (
            f"{int(differ.sum())} of {differ.size} values differ, max relative "
            f"{float((np.abs(actual - expected) / np.maximum(np.abs(expected), 1e-30)).max()):.3g}")
        return
    np.testing.assert_allclose(actual, expected, rtol=1e-5, atol=1e-7 * float(np.abs(expected).max()))
