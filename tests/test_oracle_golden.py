"""The oracle against the reference's closed forms and known answers.

Pins the C restatement (oracle/sg_oracle.c) before it is trusted as the
checker of the HIP path.
"""
import math

import numpy as np
import pytest

import cases
import oracle


def make_oracle(dimensions, options, catalogue=None, sky_cells=None, milky_way=None):
    return oracle.OracleTracer(dimensions, options, "f32", catalogue=catalogue, sky_cells=sky_cells,
                               milky_way=milky_way)


@pytest.mark.parametrize("case", cases.ALL, ids=lambda c: c.__name__)
def test_oracle_reference_case(case):
    case(make_oracle)


def test_rng_known_answers():
    """lowbias32 hash and PCG-RXS-M-XS step, restated from their published
    definitions (and src/shaders/pathtrace.wgsl:87-98)."""
    lib = oracle.library("f32")

    def hash_py(x):
        x = ((x ^ (x >> 16)) * 0x7FEB352D) & 0xFFFFFFFF
        x = ((x ^ (x >> 15)) * 0x846CA68B) & 0xFFFFFFFF
        return x ^ (x >> 16)

    for v in (0, 1, 2, 0x9E3779B9, 0xDEADBEEF, 0xFFFFFFFF, 123456789):
        assert lib.orc_hash(v) == hash_py(v)
    # PCG32 RXS-M-XS: the step constants are the LCG of O'Neill's paper.
    import ctypes

    s = ctypes.c_uint32(42)
    state = 42
    for _ in range(1000):
        r = lib.orc_random(ctypes.byref(s))
        state = (state * 747796405 + 2891336453) & 0xFFFFFFFF
        word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
        expect = (((word >> 22) ^ word) >> 8) / 16777216.0
        assert r == expect and 0.0 <= r < 1.0
    assert s.value == state
    # lowbias32 is a bijection: no collisions on a block of inputs.
    vals = {lib.orc_hash(v) for v in range(1 << 16)}
    assert len(vals) == 1 << 16


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_ring_energy_at_most_one(precision):
    """src/ring_tests.rs:351-412: transmitted + scattered energy <= 1."""
    lib = oracle.library(precision)
    for albedo in (0.8, 0.88, 1.0):
        for tau in (0.001, 0.1, 0.7, 2.0):
            for mu in (0.05, 0.5, 1.0):
                e = lib.orc_ring_energy(albedo, tau, mu, 4096)
                assert 0.0 <= e <= 1.0001, (albedo, tau, mu, e)
    # White, thin slab: nothing absorbed, so energy is ~1 (conservation, not just a bound).
    assert abs(lib.orc_ring_energy(1.0, 0.001, 1.0, 4096) - 1.0) < 2e-3


def test_fp64_oracle_builds_are_distinct():
    assert oracle.library("f32").orc_real_bytes() == 4
    assert oracle.library("f64").orc_real_bytes() == 8


def test_oracle_constant_ring_table_is_the_unbanded_ring():
    """The oracle's tabulated-profile lookup: a table holding the body's own
    optical depth everywhere renders the unbanded ring bit for bit, and a
    different table changes the ring."""
    import numpy as np

    import oracle
    from stargaze_amd import scenes
    from stargaze_amd.pathtracer import Options

    size = (32, 18)
    frame = scenes.ringed_giant(*size)
    b = frame.bodies.copy()
    p = b["ring_params"][1].copy()
    p[2] = 0.0
    b["ring_params"][1] = p
    frame = scenes.Frame(frame.globals, b, frame.ground)
    images = []
    for table in (None, np.full(512, float(p[1]), np.float32), np.full(512, 0.05, np.float32)):
        t = oracle.OracleTracer(size, Options(samples_per_frame=4, max_bounces=3, milky_way=0.0))
        if table is not None:
            t.set_ring_profile(1, table)
        t.encode(frame)
        images.append(t.read_accumulation())
    assert np.array_equal(images[0], images[1])
    assert not np.array_equal(images[0], images[2])


def test_oracle_eclipse_energy_against_unthinned_planetshine():
    """src/motion_noise_tests.rs:361-405 on the oracle: production against
    the never-thinned estimator of the same oracle. In the planetshine view
    the eclipsed air sees only the moon's light, thinned once samples
    accumulate (the two means differ by 2e-4); the sun view has no
    reflector, so there the estimates are equal."""
    def reference(dimensions, options):
        t = oracle.OracleTracer(dimensions, options, "f32")
        t.set_planetshine_thinning(False)
        return t

    out = cases.eclipse_energy(lambda d, o: make_oracle(d, o), reference)
    print(out)
    assert out["planetshine"][0] != out["planetshine"][1], "the thinning did not act"


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_oracle_value_noise_is_continuous(precision):
    t = oracle.OracleTracer((1, 1), cases.opts(), precision)
    from stargaze_amd import scenes
    from stargaze_amd.pathtracer import PROBE_NOISE

    jumps, spans = cases.noise_continuity(lambda rays: t.probe(scenes.frame(), PROBE_NOISE, rays))
    print(precision, jumps, spans)
