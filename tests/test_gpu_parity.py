"""HIP path against the oracle on MI355X: per-primitive and per-image parity."""
import math

import numpy as np
import pytest

import cases
import oracle
import parity
from stargaze_amd import pathtracer as ptm
from stargaze_amd import scenes
from stargaze_amd.layouts import MISS, PROBE_RAY
from stargaze_amd.pathtracer import Options, PathTracer

pytestmark = pytest.mark.gpu


def make_gpu(dimensions, options, catalogue=None, sky_cells=None, milky_way=None):
    t = PathTracer(dimensions, catalogue=catalogue, sky_cells=sky_cells, options=options)
    if milky_way is not None:
        t.set_milky_way(milky_way)
    return t


@pytest.mark.parametrize("case", cases.ALL, ids=lambda c: c.__name__)
def test_reference_case_on_gpu(case):
    print(f"case {case.__name__}", flush=True)
    case(make_gpu)


def _rays_at_bodies(frame, rng, n_per_target, anchors):
    """Rays from each anchor (camera or a body surface) aimed inside each
    other body's disc (within 0.9 of its angular radius)."""
    b = frame.bodies
    centres = b["center"].astype(np.float64) + b["center_low"][:, :3].astype(np.float64)
    radii = b["radius"].astype(np.float64)
    rays = []
    for anchor in anchors:
        for target in range(len(b)):
            if anchor == target:
                continue
            for _ in range(n_per_target):
                if anchor == MISS:
                    origin = np.zeros(3)
                    offset = np.zeros(3)
                else:
                    nrm = rng.normal(size=3)
                    nrm /= np.linalg.norm(nrm)
                    offset = nrm * radii[anchor] * (1.0 + 2.0e-6)
                    origin = centres[anchor] + offset
                to = centres[target] - origin
                dist = np.linalg.norm(to)
                axis = to / dist
                sin_max = min(radii[target] / dist, 1.0)
                t1 = np.cross(axis, [0.0, 0.0, 1.0] if abs(axis[2]) < 0.9 else [0.0, 1.0, 0.0])
                t1 /= np.linalg.norm(t1)
                t2 = np.cross(axis, t1)
                r = 0.9 * sin_max * math.sqrt(rng.uniform())
                phi = rng.uniform(0, 2 * math.pi)
                d = axis + r * (math.cos(phi) * t1 + math.sin(phi) * t2)
                d /= np.linalg.norm(d)
                ray = np.zeros((), PROBE_RAY)
                ray["offset"] = offset
                ray["anchor"] = anchor
                ray["direction"] = d
                ray["target"] = target
                rays.append(ray)
    return np.array(rays, PROBE_RAY)


@pytest.mark.parametrize("scene", ["airless_triple", "ringed_giant"])
def test_sphere_hits_match_fp64_oracle(scene):
    """Per-primitive hit distances within 1e-6 relative of the FP64 oracle."""
    frame = getattr(scenes, scene)(64, 64)
    rng = np.random.default_rng(7)
    anchors = [MISS] + [i for i in range(len(frame.bodies)) if frame.bodies["emissive"][i] < 0.5]
    rays = _rays_at_bodies(frame, rng, 64, anchors)
    gpu = PathTracer((4, 4), options=Options(milky_way=0.0)).probe(frame, ptm.PROBE_SPHERE, rays)
    ref = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f64").probe(frame, ptm.PROBE_SPHERE, rays)
    assert np.all(ref["index"] != MISS), "probe rays must hit their targets"
    assert np.array_equal(gpu["index"], ref["index"].astype(np.uint32))
    rel = np.abs(gpu["distance"].astype(np.float64) - ref["distance"]) / ref["distance"]
    print(f"{scene}: {rays.size} rays, max rel {rel.max():.3e}")
    assert rel.max() <= 1e-6, rel.max()
    dot = np.einsum("ij,ij->i", gpu["normal"].astype(np.float64), ref["normal"])
    assert np.all(dot > 1 - 1e-6)


def test_ring_hits_match_fp64_oracle():
    frame = scenes.ringed_giant(64, 64)
    rng = np.random.default_rng(3)
    giant = 1
    b = frame.bodies[giant]
    centre = b["center"].astype(np.float64) + b["center_low"][:3].astype(np.float64)
    pole = b["ring_plane"][:3].astype(np.float64)
    e1 = np.cross(pole, [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(pole, e1)
    rays = np.zeros(512, PROBE_RAY)
    for i in range(rays.size):
        rr = b["radius"] * rng.uniform(b["ring_plane"][3] + 0.01, b["ring_params"][0] - 0.01)
        phi = rng.uniform(0, 2 * math.pi)
        target = centre + rr * (math.cos(phi) * e1 + math.sin(phi) * e2)
        d = target / np.linalg.norm(target)
        rays[i]["anchor"] = MISS
        rays[i]["direction"] = d
        rays[i]["target"] = giant
    gpu = PathTracer((4, 4), options=Options(milky_way=0.0)).probe(frame, ptm.PROBE_RING, rays)
    ref = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f64").probe(frame, ptm.PROBE_RING, rays)
    hit = ref["index"] != MISS
    assert hit.mean() > 0.95
    assert np.array_equal(gpu["index"][hit], ref["index"][hit].astype(np.uint32))
    rel = np.abs(gpu["distance"][hit] - ref["distance"][hit]) / ref["distance"][hit]
    print(f"ring: max rel distance {rel.max():.3e}")
    assert rel.max() <= 1e-6
    # Transmission through the banded profile: same optical depth to 1e-4
    # (the f32 profile's sin(6100 x) terms amplify radius rounding).
    t_rel = np.abs(gpu["normal"][hit, 0] - ref["normal"][hit, 0]) / np.maximum(ref["normal"][hit, 0], 1e-12)
    assert np.median(t_rel) < 1e-4


def _render_pair(frame, size, spp, bounces, batches, profile=None):
    opts = Options(samples_per_frame=spp, max_bounces=bounces, milky_way=0.0)
    gpu = PathTracer(size, options=opts)
    gpu.enable_moments(True)
    orc = oracle.OracleTracer(size, Options(samples_per_frame=spp, max_bounces=bounces, milky_way=0.0), "f32")
    orc.enable_moments(True)
    if profile is not None:
        gpu.set_ring_profile(*profile)
        orc.set_ring_profile(*profile)
    for _ in range(batches):
        gpu.encode(frame)
        orc.encode(frame)
    return (gpu.read_accumulation(), gpu.read_moments()), (orc.read_accumulation(), orc.read_moments())


@pytest.mark.parametrize(
    "scene,size,spp,bounces,batches",
    [
        ("airless_triple", (64, 64), 16, 4, 2),
        ("ringed_giant", (96, 54), 16, 6, 2),
    ],
)
def test_images_match_oracle_per_tile(scene, size, spp, bounces, batches):
    frame = getattr(scenes, scene)(*size)
    (ga, gm), (oa, om) = _render_pair(frame, size, spp, bounces, batches)
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    agree = parity.pixel_agreement(ga, oa)
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    print(f"{scene}: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.99, agree


def test_tabulated_ring_matches_oracle_per_tile():
    """The ringed giant with its 512-bin C/B/A optical-depth table."""
    size = (96, 54)
    frame = scenes.ringed_giant(*size)
    profile = (1, scenes.ring_profile_cba(frame, 1))
    (ga, gm), (oa, om) = _render_pair(frame, size, 16, 6, 2, profile)
    agree = parity.pixel_agreement(ga, oa)
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    print(f"tabulated ring: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.99, agree


def test_constant_ring_table_is_the_unbanded_ring():
    """A table holding the body's own optical depth everywhere renders the
    unbanded ring bit for bit."""
    size = (64, 36)
    frame = scenes.ringed_giant(*size)
    b = frame.bodies.copy()
    p = b["ring_params"][1].copy()
    p[2] = 0.0  # unbanded
    b["ring_params"][1] = p
    frame = scenes.Frame(frame.globals, b, frame.ground)
    images = []
    for table in (None, np.full(512, float(p[1]), np.float32)):
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=4, milky_way=0.0))
        if table is not None:
            t.set_ring_profile(1, table)
        t.encode(frame)
        images.append(t.read_accumulation())
        t.close()
    assert np.array_equal(images[0], images[1])


def test_batch_size_invariance_with_air():
    """src/motion_noise_tests.rs:326-363: with air on, 1 spp x N and a batch
    of N give bitwise-identical accumulations."""
    f = cases.air_frame(cases.earthlike())
    f.bodies = np.concatenate([f.bodies, scenes.bodies(scenes.body([1.0, 0.0, -1.0], 0.005, [100000.0] * 3, True))])
    f.globals["viewport"] = [16, 9, 1, 0]
    f.globals["cam_forward"][3] = 0.5
    single = PathTracer((16, 9), options=Options(max_bounces=1, sample_limit=19, milky_way=0.0))
    batched = PathTracer((16, 9), options=Options(max_bounces=1, sample_limit=19, samples_per_frame=64, milky_way=0.0))
    a = cases.sample(single, f, 19)
    b = cases.sample(batched, f, 2)
    assert single.samples() == 19 and batched.samples() == 19
    assert np.array_equal(a, b)


def test_seed_determinism_and_reset():
    """src/pathtracer_tests.rs:199-237: same seed, same image; other seed or a
    reset gives a different noise sequence."""
    f = scenes.frame()
    f.globals["viewport"] = [32, 32, 1, 1]
    f.globals["cam_forward"][3] = 0.5
    f.bodies = scenes.bodies(scenes.body([0, 0, -3], 1.0, [1.0] * 3, True), scenes.body([0, 2, -3], 1.0, [0.5] * 3, False))

    def make(seed):
        return PathTracer((32, 32), options=Options(noise_seed=seed, milky_way=0.0))

    t = make(0)
    first = cases.sample(t, f, 1)
    f.scene_time = 1.0
    reset = cases.sample(t, f, 1)
    assert t.samples() == 1 and not np.array_equal(first, reset)
    f.scene_time = 0.0
    assert np.array_equal(first, cases.sample(make(0), f, 1))
    assert not np.array_equal(first, cases.sample(make(12345), f, 1))


@pytest.mark.parametrize("radius", [6_371_000.0, 5_000.0])
def test_terrain_probe_matches_exhaustive_triangles(radius):
    import terrain_cases as tc

    cells = 64
    heights, data = tc.make_terrain(cells)
    rays = tc.probe_rays(radius, extra=200)
    t = PathTracer((4, 4), options=Options(milky_way=0.0))
    t.upload_terrain(data, cells)
    res = t.probe(tc.probe_frame(cells, radius), ptm.PROBE_TERRAIN, rays)
    tc.check(res, tc.exhaustive(heights, cells, radius, rays))


def _render_halo(size, spp, bounces, batches):
    frame, data, cells = scenes.halo_parity(*size)
    opts = dict(samples_per_frame=spp, max_bounces=bounces, milky_way=0.0)
    gpu = PathTracer(size, options=Options(**opts))
    gpu.upload_terrain(data, cells)
    gpu.enable_moments(True)
    orc = oracle.OracleTracer(size, Options(**opts), "f32")
    orc.upload_terrain(data, cells)
    orc.enable_moments(True)
    for _ in range(batches):
        gpu.encode(frame)
        orc.encode(frame)
    return frame, (gpu.read_accumulation(), gpu.read_moments()), (orc.read_accumulation(), orc.read_moments())


def test_halo_scene_matches_oracle_per_tile():
    """Atmosphere, sky cache, terrain, props and two suns at reference capacity."""
    frame, (ga, gm), (oa, om) = _render_halo((96, 54), 8, 4, 2)
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    agree = parity.pixel_agreement(ga, oa)
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    print(f"halo: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}, "
          f"mean {ga[..., :3].mean():.4g} vs {oa[..., :3].mean():.4g}")
    assert agree >= 0.99, agree


def test_landscape_probe_matches_oracle():
    """Closest landscape hits (host sphere in metres, terrain, props) from the
    camera, against the FP64 oracle."""
    frame, data, cells = scenes.halo_parity(32, 18)
    rng = np.random.default_rng(11)
    n = 2048
    rays = np.zeros(n, PROBE_RAY)
    dirs = rng.normal(size=(n, 3))
    dirs[:, 1] = -np.abs(dirs[:, 1]) * 0.3 - 0.02
    up = frame.globals["ground_up"][:3].astype(np.float64)
    east = frame.globals["ground_east"][:3].astype(np.float64)
    north = frame.globals["ground_north"][:3].astype(np.float64)
    world = dirs[:, :1] * east + dirs[:, 1:2] * up + dirs[:, 2:3] * north
    world /= np.linalg.norm(world, axis=1, keepdims=True)
    rays["anchor"] = MISS
    rays["direction"] = world
    gpu = PathTracer((4, 4), options=Options(milky_way=0.0))
    gpu.upload_terrain(data, cells)
    g = gpu.probe(frame, ptm.PROBE_LANDSCAPE, rays)
    orc = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f64")
    orc.upload_terrain(data, cells)
    r = orc.probe(frame, ptm.PROBE_LANDSCAPE, rays)
    same = g["index"] == r["index"].astype(np.uint32)
    # Host sphere vs coincident terrain near the levelled lookout is a tie
    # broken differently by the FP64 and FP32 quadratics (distances agree).
    tie = ~same & np.isin(g["index"], [0x80000000, 0x80100000]) & np.isin(r["index"], [0x80000000, 0x80100000])
    assert np.all(same | tie)
    tie_rel = np.abs(g["distance"][tie] - r["distance"][tie]) / r["distance"][tie]
    assert np.all(tie_rel < 1e-5)
    hit = same & (r["index"] != MISS)
    rel = np.abs(g["distance"][hit] - r["distance"][hit]) / r["distance"][hit]
    print(f"landscape: {hit.sum()} hits, same index {same.mean():.4f}, median rel {np.median(rel):.2e}, max {rel.max():.2e}")
    # Host sphere (FP64 on both sides), terrain triangles and boxes and
    # ellipsoids in FP32: the median is at the FP32 rounding of the distance,
    # the worst case a grazing terrain triangle (measured 1.9e-6).
    assert np.median(rel) <= 1e-6 and rel.max() <= 1e-5


@pytest.mark.parametrize("scene", ["airless_triple", "ringed_giant", "halo"])
def test_wavefront_and_fused_integrators_agree(scene):
    """Both integrators run the same step functions in the same draw order.

    The images are bit-identical, except on the ground scene with air, whose
    deferred camera-air shadow rays the wavefront integrator sums on their
    own (parity.assert_same_draws). Chunking the wavefront differently
    changes nothing: those images are bit-identical.
    """
    size = (48, 27)
    if scene == "halo":
        frame, data, cells = scenes.halo_parity(*size)
    else:
        frame, data, cells = getattr(scenes, scene)(*size), None, None
    images = []
    for integrator, chunk in ((ptm.INTEGRATOR_FUSED, 0), (ptm.INTEGRATOR_WAVEFRONT, 0), (ptm.INTEGRATOR_WAVEFRONT, 1000)):
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=5, milky_way=0.0))
        if data is not None:
            t.upload_terrain(data, cells)
        t.set_integrator(integrator, chunk)
        t.encode(frame)
        t.encode(frame)
        images.append(t.read_accumulation())
    parity.assert_same_draws(images[1], images[0], exact=scene != "halo")
    assert np.array_equal(images[1], images[2])


def test_deferred_air_queue_overflow_reruns_the_chunk():
    """Camera-air shadow rays whose landscape query the wavefront integrator
    defers to a kernel of its own: a queue too small for a chunk's pushes
    grows and the chunk runs again, so the image is bit-identical to one
    rendered with room to spare, and agrees with the fused integrator's,
    which answers every query inline (to the last bits of the sums)."""
    size = (48, 27)
    frame, data, cells = scenes.halo_parity(*size)
    images, queues = [], []
    for integrator, queue in ((ptm.INTEGRATOR_FUSED, None), (ptm.INTEGRATOR_WAVEFRONT, 64),
                              (ptm.INTEGRATOR_WAVEFRONT, 1)):
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=3, milky_way=0.0))
        t.upload_terrain(data, cells)
        t.set_integrator(integrator, 0)
        if queue is not None:
            t.set_air_queue(queue)
        t.encode(frame)
        t.encode(frame)
        images.append(t.read_accumulation())
        queues.append(t.air_queue)
        t.close()
    assert queues[1] == 64 and queues[2] > 1, queues
    parity.assert_same_draws(images[1], images[0], exact=False)
    assert np.array_equal(images[1], images[2])


def test_step_times_cover_the_wavefront_kernels():
    """sg_step_time: while profiling, each wavefront step reports device time
    and launches. Camera air runs once per chunk (its first round), and on a
    ground scene with air so do the deferred shadow-ray kernels; a frame
    without air runs neither. The steps fit in the pass's kernel time."""
    size = (64, 36)
    frame, data, cells = scenes.halo_parity(*size)
    steps = (ptm.STEP_EXTEND, ptm.STEP_CAMERA_AIR, ptm.STEP_SHADE, ptm.STEP_AIR_OCCLUSION)
    for scene, ground in (("halo", True), ("airless_triple", False)):
        f = frame if ground else scenes.airless_triple(*size)
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=4, milky_way=0.0))
        if ground:
            t.upload_terrain(data, cells)
        t.set_integrator(ptm.INTEGRATOR_WAVEFRONT, 4096)
        t.set_profiling(True)
        t.encode(f)
        t.encode(f)
        t.synchronize()
        chunks = 2 * -(-(size[0] * size[1]) // (4096 // 8))
        times = {s: t.step_time(s) for s in steps}
        kernel_ms, _ = t.kernel_time()
        t.close()
        print(scene, times, f"kernel {kernel_ms:.3f} ms")
        assert times[ptm.STEP_EXTEND][1] >= chunks and times[ptm.STEP_SHADE][1] == times[ptm.STEP_EXTEND][1]
        expected = chunks if ground else 0
        assert times[ptm.STEP_CAMERA_AIR][1] == expected and times[ptm.STEP_AIR_OCCLUSION][1] == expected
        assert all((ms > 0.0) == (n > 0) for ms, n in times.values())
        assert sum(ms for ms, _ in times.values()) <= kernel_ms * 1.01


@pytest.mark.parametrize("scene", ["ringed_giant", "halo"])
def test_wavefront_tail_and_autotune(scene):
    """The fused tail below the compaction threshold changes no draw: images
    agree with the plain wavefront's; the autotuner's choice renders the
    same image as well."""
    size = (64, 36)
    if scene == "halo":
        frame, data, cells = scenes.halo_parity(*size)
    else:
        frame, data, cells = getattr(scenes, scene)(*size), None, None

    def render(configure):
        t = PathTracer(size, options=Options(samples_per_frame=8, max_bounces=6, milky_way=0.0))
        if data is not None:
            t.upload_terrain(data, cells)
        configure(t)
        t.encode(frame)
        t.encode(frame)
        img = t.read_accumulation()
        t.close()
        return img

    base = render(lambda t: t.set_integrator(ptm.INTEGRATOR_WAVEFRONT, 4096))
    for tail in (64, 1024, 1 << 20):
        img = render(lambda t: (t.set_integrator(ptm.INTEGRATOR_WAVEFRONT, 4096), t.set_tail(tail)))
        parity.assert_same_draws(img, base, exact=scene != "halo")

    def tuned(t):
        best, ms = t.autotune(frame, repeats=1)
        print(f"{scene}: autotune {best} {ms:.3f} ms")
        assert best == t.launch and ms > 0.0
        assert t.samples() == 0

    img = render(tuned)
    parity.assert_same_draws(img, base, exact=scene != "halo")


def test_eclipse_energy_against_the_oracle():
    """src/motion_noise_tests.rs:361-405: 4096 spp of the HIP production
    estimator against 4096 spp of the oracle's never-thinned estimator, on
    the eclipsed-planetshine and eclipsed-sun views, within 3% in mean
    luminance."""
    def reference(dimensions, options):
        t = oracle.OracleTracer(dimensions, options, "f32")
        t.set_planetshine_thinning(False)
        return t

    out = cases.eclipse_energy(make_gpu, reference)
    for name, (actual, expected) in out.items():
        print(f"{name}: gpu {actual:.6e} oracle {expected:.6e} ratio {actual / expected:.5f}")


def test_value_noise_is_continuous_and_matches_oracle():
    """src/noise_tests.rs:5-73 restated on the HIP noise (SG_PROBE_NOISE),
    and the HIP hash, value noise and 6-octave fbm against the FP32 oracle
    point by point."""
    gpu = PathTracer((1, 1), options=Options(milky_way=0.0))
    probe = lambda rays: gpu.probe(scenes.frame(), ptm.PROBE_NOISE, rays)  # noqa: E731
    jumps, spans = cases.noise_continuity(probe)
    rays, _, _ = cases.noise_points()
    got = probe(rays)
    want = oracle.OracleTracer((1, 1), Options(milky_way=0.0), "f32").probe(scenes.frame(), ptm.PROBE_NOISE, rays)
    dv = np.abs(got["distance"] - want["distance"]).max()
    dh = np.abs(got["normal"][:, 0] - want["normal"][:, 0]).max()
    df = np.abs(got["normal"][:, 1] - want["normal"][:, 1]).max()
    print(f"noise: worst plane jump {jumps:.2e}, patch spans {spans}; vs oracle: vnoise {dv:.2e}, "
          f"hash {dh:.2e}, fbm {df:.2e}")
    assert dv <= 1e-6 and dh <= 1e-6 and df <= 1e-6
    gpu.close()


@pytest.mark.parametrize("theme", [0, 1], ids=["forest", "desert rocks"])
def test_prop_kinds_match_oracle(theme):
    """Camera rays at every primitive of the reference's lookout scenery
    (boxes, stone rocks, tumbled rocks, conifers, grass, logs) through the
    HIP prop BVH and terrain, against the FP32 oracle's flat loop and the
    FP64 build: the same element, median relative distance within 1e-6, and
    the 99th percentile within what FP32 reaches per kind (as
    tests/test_ground.py records it for the FP32 oracle against FP64)."""
    import ground_cases as gc

    f, data, cells = scenes.lookout(48, 27, theme=theme)
    rays, kinds = gc.prop_rays(f)
    gpu = PathTracer((4, 4), options=Options(milky_way=0.0))
    gpu.upload_terrain(data, cells)
    g = gpu.probe(f, ptm.PROBE_LANDSCAPE, rays)
    gpu.close()
    o32 = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f32")
    o32.upload_terrain(data, cells)
    o64 = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f64")
    o64.upload_terrain(data, cells)
    b, a = o32.probe(f, ptm.PROBE_LANDSCAPE, rays), o64.probe(f, ptm.PROBE_LANDSCAPE, rays)
    p99_fp64 = {0.0: 1e-6, 2.0: 1e-6, 2.25: 1e-6, 4.0: 1e-5, 5.0: 1e-5, 3.0: 1e-3}
    for k in np.unique(kinds):
        sel = kinds == k
        name = gc.KIND_NAMES[k]
        same32 = g["index"][sel] == b["index"][sel].astype(np.uint32)
        same64 = g["index"][sel] == a["index"][sel].astype(np.uint32)
        hit32 = same32 & (g["index"][sel] != MISS)
        hit64 = same64 & (g["index"][sel] != MISS)
        r32 = np.abs(g["distance"][sel][hit32] - b["distance"][sel][hit32]) / b["distance"][sel][hit32]
        r64 = np.abs(g["distance"][sel][hit64] - a["distance"][sel][hit64]) / a["distance"][sel][hit64]
        print(f"{name}: {sel.sum()} rays; vs fp32 same {same32.mean():.4f} p99 {np.percentile(r32, 99):.2e} "
              f"max {r32.max():.2e}; vs fp64 same {same64.mean():.4f} median {np.median(r64):.2e} "
              f"p99 {np.percentile(r64, 99):.2e} max {r64.max():.2e}")
        assert same32.mean() >= 0.995 and same64.mean() >= 0.99, name
        assert np.median(r32) <= 1e-6 and np.percentile(r32, 99) <= p99_fp64[k], name
        assert np.median(r64) <= 1e-6 and np.percentile(r64, 99) <= p99_fp64[k], name


@pytest.mark.parametrize("theme", [0, 1], ids=["forest", "desert rocks"])
def test_lookout_matches_oracle_per_tile(theme):
    """The lookout scenery rendered with air, sky cache and one sun: every
    primitive kind's shape and material (stone, foliage, bark, grass) against
    the FP32 oracle per 16x16 tile."""
    size = (96, 54)
    f, data, cells = scenes.lookout(*size, theme=theme)
    opts = dict(samples_per_frame=8, max_bounces=4, milky_way=0.0)
    gpu = PathTracer(size, options=Options(**opts))
    gpu.upload_terrain(data, cells)
    gpu.enable_moments(True)
    orc = oracle.OracleTracer(size, Options(**opts), "f32")
    orc.upload_terrain(data, cells)
    orc.enable_moments(True)
    for _ in range(2):
        gpu.encode(f)
        orc.encode(f)
    ga, gm, oa, om = gpu.read_accumulation(), gpu.read_moments(), orc.read_accumulation(), orc.read_moments()
    assert np.all(np.isfinite(ga)) and np.all(ga[..., :3] >= 0)
    agree = parity.pixel_agreement(ga, oa)
    z = parity.tile_z(ga, gm, oa, om)
    over, worst = parity.assert_tiles(z)
    print(f"lookout theme {theme}: pixels agreeing {agree:.4f}, tiles over 3 sigma {over:.4f}, worst z {worst:.2f}")
    assert agree >= 0.99, agree
    gpu.close()
