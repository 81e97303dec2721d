"""The allocation-free display entry point (sg_display_srgb8_noalloc)
against sg_display_srgb8, through the C ABI on the GPU."""
import numpy as np
import pytest

from stargaze_amd import scenes
from stargaze_amd._native import NativeError
from stargaze_amd.pathtracer import Options, PathTracer

pytestmark = pytest.mark.gpu


def test_noalloc_display_matches_allocating_display_twice():
    print("no-alloc display", flush=True)
    size = (96, 64)
    t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=3, milky_way=0.0))
    t.encode(scenes.airless_triple(*size))
    # Two calls in a row, the second smaller: each equals the allocating path.
    for w, h in ((320, 180), (37, 23)):
        expected = t.display_srgb8(w, h)
        got = t.display_srgb8_noalloc(w, h)
        assert got.shape == (h, w, 4)
        assert np.array_equal(got, expected), (w, h, int(np.abs(got.astype(int) - expected).max()))
        assert np.any(got[..., :3] != 0)
    t.close()


def test_noalloc_display_rejects_targets_beyond_its_staging():
    t = PathTracer((8, 8), options=Options(milky_way=0.0))
    with pytest.raises(NativeError, match="capacity"):
        t.display_srgb8_noalloc(3840, 2161)
    t.close()
