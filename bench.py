"""Benchmark: progressive path-tracing passes on one MI355X.

A step is one progressive pass (``PathTracer::encode``) of ``spp`` path
samples per pixel over the workload image, with the scene already resident
on the GPU. ``value`` is path samples per second over the timed steps.

Prints one JSON line. ``roofline`` prices the integrator kernel's algorithmic
work (event counts of the CPU oracle on a band of the same workload times
the FLOP cost table below) against its HIP-event time; ``cpu_baseline`` times
the FP32 oracle on the host cores on a bounded band of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# The oracle module lives in oracle/oracle.py; put its directory first so the
# directory itself is not taken for a namespace package.
sys.path.insert(0, str(ROOT / "oracle"))

from stargaze_amd import scenes  # noqa: E402
from stargaze_amd.pathtracer import Options, PathTracer  # noqa: E402

METRIC = "converged Mpath-samples/s at 4K Halo @ fixed RMSE; adjoint:forward time ratio"

# FP32 operations per event of the work model (transcendentals count 4).
# Derived from the operation counts of the corresponding WGSL functions.
FLOP_COST = {
    "paths": 30,          # camera_direction, seeding, accumulation share
    "segments": 10,       # closest-hit / ring loop bookkeeping per segment
    "sphere": 40,         # relative_center + cross-product discriminant
    "sphere_hit": 60,     # roots, distance selection, body-local normal
    "ring": 45,           # plane hit via perpendicular decomposition
    "light_sample": 90,   # cone width, two draws, basis_direction
    "visibility": 5,      # shadow-ray bookkeeping (its tests are counted)
    "ring_optics": 45,    # banded optical depth, four sines
    "atmo_cell": 45,      # density, extinction, segment integral
    "density_step": 14,   # one light-path density sample
    "vnoise": 195,        # eight hash31 lattice values + trilinear mix
    "box": 30,
    "rock": 150,
    "vegetation": 400,
    "terrain_node": 25,
    "triangle": 35,
    "vertex": 70,         # surface shading: sampling frame, throughput
    "rng": 0,             # integer work
}
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32, dense
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the camera-air kernel on the Halo 4K pass (2^24
# paths): rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE (the factors of 1 GiB copies),
# in KiB summed over the pass's 16 launches; not measured by this script,
# since the counters need a profiler run of their own.
TRAFFIC_CAMERA_AIR_PER_LAUNCH = {"halo_4k": (19014888 * 2 + 49072105) * 1024 / 16}
TRAFFIC_SOURCE = "profiles/halo_4k_kernels.md (rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per k_camera_air launch)"
# Path state the camera-air step must read and write per path: ray origin
# (FP64) and direction, the hit record, throughput, radiance and the RNG.
CAMERA_AIR_STATE_BYTES = 32 + 16 + 48 + 2 * 16 + 2 * 16 + 2 * 4

def _halo(width, height):
    frame, data, cells = scenes.halo_observatory(width, height)
    return frame, (data, cells), None, scenes.HALO_AIR


def _giant(width, height):
    """The ringed giant with the 512-bin C/B/A optical-depth table."""
    frame = scenes.ringed_giant(width, height)
    return frame, None, (1, scenes.ring_profile_cba(frame, 1)), None


def _plain(builder):
    return lambda width, height: (builder(width, height), None, None, None)


WORKLOADS = {
    # name: (builder -> (frame, terrain, ring profile, air model), width, height, spp per pass, max bounces)
    "halo_4k": (_halo, 3840, 2160, 32, 4),
    "ringed_giant_1080p": (_giant, 1920, 1080, 64, 6),
    "airless_triple_256": (_plain(scenes.airless_triple), 256, 256, 16, 4),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def oracle_tracer(width, height, spp, bounces, terrain, profile=None, air=None):
    import oracle

    t = oracle.OracleTracer((width, height), Options(samples_per_frame=spp, max_bounces=bounces, milky_way=0.0))
    if terrain is not None:
        t.upload_terrain(*terrain)
    if profile is not None:
        t.set_ring_profile(*profile)
    if air is not None:
        t.set_air_model(**air)
    return t


PROP_EVENTS = ("box", "rock", "vegetation")
# Terrain work of the camera-air step's shadow rays: the wavefront integrator
# runs it in k_air_occlusion, not in k_camera_air.
LANDSCAPE_EVENTS = ("terrain_node", "triangle")


def work_model(frame, terrain, profile, air, width, height, spp, bounces, rows=16):
    """Event counts per path sample of a steady-state pass (not the first,
    which takes extra light strata near limbs) from the oracle on four bands
    of rows (1/8, 3/8, 5/8, 7/8 of the height), for the whole pass and for
    its camera-air step. FLOP are the counts times FLOP_COST, without the
    primitive tests of the oracle's flat skip-count loop (the GPU traverses a
    BVH over the primitives instead, so those counts are not its work; they
    are reported apart)."""
    import oracle

    t = oracle_tracer(width, height, spp, bounces, terrain, profile, air)
    t.encode(frame, rows=(0, 1))  # first pass: fills the sky and terrain caches
    rest, cam = None, None
    bands = [height * (2 * k + 1) // 8 for k in range(4)]
    per_band = max(1, rows // len(bands))
    for y0 in bands:
        t.history_samples = spp
        a, b = oracle.count_events(t, frame, (y0, y0 + per_band), stages=True)
        rest = a if rest is None else {k: rest[k] + a[k] for k in rest}
        cam = b if cam is None else {k: cam[k] + b[k] for k in cam}
    n = width * per_band * len(bands) * spp

    def price(ev):
        per = {k: v / n for k, v in ev.items()}
        flops = sum(per[k] * FLOP_COST[k] for k in FLOP_COST if k not in PROP_EVENTS)
        props = sum(per[k] * FLOP_COST[k] for k in PROP_EVENTS)
        return per, flops, props

    whole = {k: rest[k] + cam[k] for k in rest}
    per, flops, props = price(whole)
    cam_per, flops_air, props_air = price(cam)
    landscape = sum(cam_per[k] * FLOP_COST[k] for k in LANDSCAPE_EVENTS)
    return {"per_event": per, "flop": flops, "flop_flat_loop_props": props,
            "flop_camera_air": flops_air - landscape, "flop_air_occlusion": landscape,
            "flop_camera_air_flat_loop_props": props_air}


def roofline(workload, wm, samples_per_step, kernel_s, pass_tflops, compulsory, air_ms, air_launches, launch,
             occlusion=(0.0, 0)):
    """FP32 vector roofline of the dominant kernel (nothing on the path is a
    dense contraction, so no MFMA instruction is issued): the camera-air
    kernel of the wavefront integrator when the frame has air, else the
    whole pass. `achieved` is the kernel's algorithmic FLOP (oracle event
    counts of the camera-air step per path sample, less the terrain work
    of its shadow rays, which k_air_occlusion does, x the paths of one
    launch) over its average launch time from HIP events on the tracer's
    stream."""
    whole = {"kernel": "all kernels of one pass", "ms": round(kernel_s * 1e3, 3),
             "flop_per_path_sample": round(wm["flop"], 1), "achieved_tflops": round(pass_tflops, 3),
             "frac": round(pass_tflops / PEAK_FP32_TFLOPS, 4),
             "flat_loop_prop_flop_per_path_sample_not_counted": round(wm["flop_flat_loop_props"], 1),
             "compulsory_bytes": compulsory}
    if air_launches == 0:
        return {"bound": "valu", "pipe": "FP32 VALU, no MFMA instructions", "achieved": round(pass_tflops, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(pass_tflops / PEAK_FP32_TFLOPS, 4),
                "traffic": None, "kernel": whole["kernel"], "pass": whole}
    per_launch_ms = air_ms / air_launches
    paths_per_launch = min(samples_per_step, launch.get("chunk_paths", samples_per_step))
    launches_per_step = math.ceil(samples_per_step / paths_per_launch)
    step_ms = air_ms / air_launches * launches_per_step
    achieved = wm["flop_camera_air"] * samples_per_step / (step_ms / 1e3) / 1e12
    return {
        "bound": "valu",
        "pipe": "FP32 VALU, no MFMA instructions",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": TRAFFIC_CAMERA_AIR_PER_LAUNCH.get(workload),
        "traffic_source": TRAFFIC_SOURCE if workload in TRAFFIC_CAMERA_AIR_PER_LAUNCH else None,
        "kernel": "k_camera_air (camera-air step of the wavefront integrator)",
        "kernel_ms_per_launch": round(per_launch_ms, 3),
        "air_occlusion": {
            "kernel": "k_air_occlusion + k_air_resolve (deferred landscape predicates of the shadow rays)",
            "ms_per_launch": round(occlusion[0] / occlusion[1], 3) if occlusion[1] else None,
            "flop_per_path_sample": round(wm.get("flop_air_occlusion", float("nan")), 1),
        },
        "paths_per_launch": paths_per_launch,
        "flop_per_path_sample": round(wm["flop_camera_air"], 1),
        "flat_loop_prop_flop_per_path_sample_not_counted": round(wm["flop_camera_air_flat_loop_props"], 1),
        "algorithmic_bytes_per_launch": paths_per_launch * CAMERA_AIR_STATE_BYTES,
        "hbm_frac": round(paths_per_launch * CAMERA_AIR_STATE_BYTES / (per_launch_ms / 1e3) / 1e9 / PEAK_HBM_GBS,
                          5),
        "pass": whole,
    }


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frame, terrain, profile, air, width, height, spp, bounces, target_s=15.0):
    """FP32 oracle on all host cores over four bands of rows at the image's
    eighth, three-eighths, five-eighths and seven-eighths heights (the work
    model's bands), sized so the sample takes ~target_s."""
    rows = 2
    while True:
        t = oracle_tracer(width, height, spp, bounces, terrain, profile, air)
        bands = [min(height - rows, height * (2 * k + 1) // 8) for k in range(4)]
        t.encode(frame, rows=(0, 1))  # fills the sky and terrain caches
        t.history_samples = spp
        t0 = time.perf_counter()
        for y0 in bands:
            t.encode(frame, rows=(y0, y0 + rows))
            t.history_samples = spp
        dt = time.perf_counter() - t0
        if dt >= 10.0 or 4 * rows >= height:
            break
        rows = min(height // 4, max(rows * 2, int(rows * target_s / max(dt, 1e-3))))
    samples = width * 4 * rows * spp
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return {
        "value": samples / dt / 1e6,
        "unit": "Mpath-samples/s",
        "cores": cores,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"FP32 oracle (C restatement of the WGSL), 4 bands of {rows} rows at 1/8, 3/8, 5/8 and 7/8 "
        f"of the height x {width} px x {spp} spp of the same workload, {dt:.1f} s",
    }


RMSE_REFERENCE = ROOT / "tests" / "golden" / "halo_rmse_reference.npz"
RMSE_META = ROOT / "tests" / "golden" / "halo_rmse_reference.json"


def rmse_leg(ms_per_step: float, spp_per_step: int, pixels_4k: int, cpu_rate: float | None) -> dict | None:
    """The fixed-RMSE half of the metric, on the Halo camera at 320 x 180.

    The reference is a 16384-spp float64-accumulated render of the same
    scene and air model (tests/golden/make_halo_reference.py); the target is
    the FP32 oracle's expected linear-RGB RMSE at 32 spp (root mean square
    over four seeds). The GPU renders the same view progressively with eight
    other seeds in passes of 8 spp; its RMSE at each spp is the root mean
    square over 16 independent seeds, and the spp where that curve crosses the target
    (log-log interpolation) is the GPU's spp to the target. At 4K those spp
    take that fraction of a 32-spp step; the CPU oracle reaches the target
    at 32 spp by construction."""
    if not RMSE_REFERENCE.exists() or not RMSE_META.exists():
        return None
    sys.path.insert(0, str(ROOT / "tests"))
    import parity

    meta = json.loads(RMSE_META.read_text())
    if "target" not in meta:
        return None
    ref = np.load(RMSE_REFERENCE)["image"]
    w, h = meta["width"], meta["height"]
    frame, data, cells = scenes.halo_observatory(w, h)
    stops = (8, 16, 24, 32, 40, 48, 64)
    # Independent draws: seeds far apart (a pass of n samples uses sequences
    # seed .. seed + n - 1) and far from the target's.
    seeds = tuple((k + 1000) * 1_000_003 for k in range(16))
    mse = np.zeros(len(stops))
    for seed in seeds:
        t = PathTracer((w, h), options=Options(noise_seed=seed, samples_per_frame=8, max_bounces=4, milky_way=0.0))
        t.upload_terrain(data, cells)
        t.set_air_model(**scenes.HALO_AIR)
        spp = 0
        for k, stop in enumerate(stops):
            while spp < stop:
                t.encode(frame)
                spp += 8
            mse[k] += parity.rmse(t.read_accumulation(), ref) ** 2 / len(seeds)
        t.close()
    curve = np.sqrt(mse)
    target = float(meta["target"]["rmse"])
    below = np.nonzero(curve <= target)[0]
    if len(below) == 0:
        spp_to_target = None
    elif below[0] == 0:
        spp_to_target = float(stops[0])
    else:
        k = int(below[0])
        x0, x1 = math.log(stops[k - 1]), math.log(stops[k])
        y0, y1 = math.log(curve[k - 1]), math.log(curve[k])
        spp_to_target = math.exp(x0 + (math.log(target) - y0) * (x1 - x0) / (y1 - y0))
    result = {
        "reference": f"{w}x{h} of the halo_4k camera, {meta['spp']} spp, float64 accumulation "
                     "(tests/golden/halo_rmse_reference.npz)",
        "target_rmse": target,
        "target_source": "FP32 oracle at 32 spp, linear RGB (W excluded), root mean square over 8 independent seeds",
        "gpu_rmse_by_spp": {str(k): round(float(v), 5) for k, v in zip(stops, curve)},
        "gpu_seeds": len(seeds),
        "gpu_spp_to_target": None if spp_to_target is None else round(spp_to_target, 2),
    }
    if spp_to_target is not None:
        gpu_s = spp_to_target / spp_per_step * ms_per_step / 1e3
        result["gpu_seconds_to_target_4k"] = round(gpu_s, 4)
        if cpu_rate:
            cpu_s = pixels_4k * 32 / (cpu_rate * 1e6)
            result["cpu_seconds_to_target_4k"] = round(cpu_s, 1)
            result["speedup_at_equal_rmse"] = round(cpu_s / gpu_s, 1)
    return result


BATCH_WORKLOADS = {
    # name: system_views keyword arguments. mixed_96 is the batched-serving
    # config (128^2 to 1080p); the others shrink the views to where per-view
    # launches leave the GPU idle.
    "mixed_96": dict(count=96, min_pixels=128 * 128, max_pixels=1920 * 1080),
    "small_96": dict(count=96, min_pixels=64 * 64, max_pixels=640 * 360),
    "tiny_96": dict(count=96, min_pixels=48 * 48, max_pixels=256 * 256),
    # Previews: small views at 1-4 spp per pass, where launches dominate.
    "preview_96": dict(count=96, min_pixels=48 * 48, max_pixels=256 * 256, spp=(1, 2, 4)),
}


def batch_leg(steps: int, warmup: int, workload: str = "mixed_96", tune: bool = True) -> dict:
    """Batched views against per-view launches on the same views.

    Both forms render identical images pass for pass (same trace code, RNG
    streams and blend order), so equal passes are equal RMSE. Latency is a
    view's completion time within a pass, measured from the pass start on
    the GPU clock (batch) or from the first launch with HIP events
    (per-view)."""
    from stargaze_amd.pathtracer import INTEGRATOR_FUSED, Batch

    t0 = time.perf_counter()
    spec = BATCH_WORKLOADS[workload]
    count = spec["count"]
    views = scenes.system_views(**spec)
    frames = [v["frame"] for v in views]
    tracers = []
    for v in views:
        t = PathTracer((v["width"], v["height"]),
                       options=Options(samples_per_frame=v["spp"], max_bounces=v["bounces"], milky_way=0.0))
        if v["terrain"] is not None:
            t.upload_terrain(*v["terrain"])
        t.set_integrator(INTEGRATOR_FUSED)
        tracers.append(t)
    batch = Batch(tracers)
    log(f"batch: {count} views built in {time.perf_counter() - t0:.1f} s")
    tuned = None
    if tune:
        best, best_ms = batch.autotune(frames, repeats=1)
        tuned = {"launch": best, "pass_ms": round(best_ms, 3)}
        log(f"batch autotune: {best} {best_ms:.2f} ms")
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    for _ in range(warmup):
        batch.encode(frames)
    batch.synchronize()
    lat_b = []
    batch.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.encode(frames)
        ms, span = batch.view_times()
        lat_b.append(ms)
    batch.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    batch_kernel_ms = batch.kernel_time()[0] / steps
    batch.set_profiling(False)
    per_view = {}
    for serial in (False, True):
        for _ in range(warmup):
            batch.encode_views(frames, serial=serial)
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            lat.append(batch.encode_views(frames, serial=serial))
        per_view[serial] = ((time.perf_counter() - t0) / steps, np.concatenate(lat))
    # Device time of the per-view kernels on one queue: the share of the
    # pass the GPU spends in them bounds what any batching can gain.
    for t in tracers:
        t.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.encode_views(frames, serial=True)
    busy_wall = (time.perf_counter() - t0) / steps
    busy_ms = sum(t.kernel_time()[0] for t in tracers) / steps
    for t in tracers:
        t.set_profiling(False)
    lat_b = np.concatenate(lat_b)
    batch.close()
    for t in tracers:
        t.close()

    def pct(a, q):
        return round(float(np.percentile(a, q)), 3)

    def form(t, lat):
        return {"views_per_s": round(count / t, 2), "pass_ms": round(t * 1e3, 3),
                "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)}}

    lo, hi = int(math.sqrt(spec["min_pixels"])), spec["max_pixels"]
    spp_set = ",".join(str(k) for k in spec.get("spp", (4, 8, 16, 32, 64)))
    return {
        "workload": workload,
        "views": count,
        "path_samples_per_pass": paths,
        "batched": form(t_batch, lat_b),
        "per_view_one_queue": form(*per_view[True]),
        "per_view_own_streams": form(*per_view[False]),
        "speedup_vs_one_queue": round(per_view[True][0] / t_batch, 3),
        "speedup_vs_own_streams": round(per_view[False][0] / t_batch, 3),
        "per_view_one_queue_kernel_ms": round(busy_ms, 3),
        "per_view_one_queue_busy_fraction": round(busy_ms / (busy_wall * 1e3), 4),
        "batched_kernel_ms": round(batch_kernel_ms, 3),
        "autotune": tuned,
        "views_desc": f"{count} random systems (PCG seed 1): log-uniform {lo}^2..{hi} px, spp {{{spp_set}}} "
        "per pass, 4 bounces; camera 2 m above a host at a random pose, 2-12 bodies, rings on half the "
        "giants, air on a third of the hosts",
    }


ADJOINT_WORKLOADS = {
    # name: (builder, width, height, spp, bounces, body with a tabulated ring, tabulated density)
    "diff_system_512": (scenes.diff_system, 512, 512, 64, 4, 2, True),
    "ringed_giant_1080p": (scenes.ringed_giant, 1920, 1080, 64, 6, 1, False),
}


def adjoint_leg(steps: int, recover_steps: int = 200) -> list:
    """Adjoint against forward kernel time on the same pass (HIP events on
    the tracer's stream). The adjoint runs each path twice (primal, then
    replay with derivatives)."""
    from stargaze_amd.pathtracer import INTEGRATOR_FUSED, INTEGRATOR_WAVEFRONT

    out = []
    for name, (builder, width, height, spp, bounces, ring_body, dense) in ADJOINT_WORKLOADS.items():
        frame = builder(width, height)
        tau = scenes.ring_profile_cba(frame, ring_body)
        density = scenes.exponential_density(frame) if dense else None

        def tracer_for(f, tau, density):
            t = PathTracer((width, height), options=Options(samples_per_frame=spp, max_bounces=bounces,
                                                            milky_way=0.0))
            t.set_ring_profile(ring_body, tau)
            if density is not None:
                t.set_density_profile(density)
            return t

        # Loss: L2 distance to a target rendered with every parameter
        # perturbed by up to 10% (PCG seed 2).
        rng = scenes.Pcg(2)
        target_frame = scenes.Frame(frame.globals.copy(), frame.bodies.copy(), frame.ground.copy())
        for b in range(len(target_frame.bodies)):
            if target_frame.bodies["emissive"][b] < 0.5:
                target_frame.bodies["color"][b] = target_frame.bodies["color"][b] * np.float32(
                    1.0 + rng.uniform(-0.1, 0.1))
        t = tracer_for(target_frame, tau * np.float32(1.0 + rng.uniform(-0.1, 0.1)),
                       None if density is None else density * np.float32(1.0 + rng.uniform(-0.1, 0.1)))
        t.encode(target_frame)
        target = t.read_accumulation()[..., :3]
        t.close()
        t = tracer_for(frame, tau, density)
        t.encode(frame)
        image = t.read_accumulation()[..., :3]
        d = (2.0 * (image - target) / image.size).astype(np.float32)
        times = {}
        # Forward passes keep per-sample primal outputs (sg_enable_adjoint):
        # the adjoint replays each path once. The rerun form (primal and
        # replay in the adjoint) is timed too.
        for integrator, label, record in ((INTEGRATOR_FUSED, "fused", False), (INTEGRATOR_FUSED, "fused", True),
                                          (INTEGRATOR_WAVEFRONT, "wavefront", True)):
            t.set_integrator(integrator)
            t.enable_adjoint(record)
            t.restart()
            t.encode(frame)
            t.adjoint(frame, d)
            t.set_profiling(True)
            for _ in range(steps):
                t.restart()
                t.encode(frame)
                t.adjoint(frame, d)
            fwd_ms, n = t.kernel_time()
            adj_ms, m = t.adjoint_time()
            times[label] = fwd_ms / max(n, 1)
            times["adjoint" if record else "adjoint_rerun"] = adj_ms / max(m, 1)
            t.set_profiling(False)
        grad = t.adjoint(frame, d)
        assert np.all(np.isfinite(grad)), "non-finite gradient"
        t.close()
        best = min(times["fused"], times["wavefront"])
        params = 3 * int(np.sum(frame.bodies["emissive"] < 0.5)) + 512 + 5 + (16 if dense else 0)
        out.append({"workload": name, "width": width, "height": height, "spp": spp, "max_bounces": bounces,
                    "parameters": params, "loss": "L2 to a target with parameters perturbed up to 10%",
                    "forward_ms": {"fused": round(times["fused"], 3), "wavefront": round(times["wavefront"], 3)},
                    "adjoint_ms": round(times["adjoint"], 3),
                    "adjoint_rerun_primal_ms": round(times["adjoint_rerun"], 3),
                    "ratio_vs_fused": round(times["adjoint"] / times["fused"], 3),
                    "ratio_vs_best_forward": round(times["adjoint"] / best, 3)})
        if name == "diff_system_512" and recover_steps > 0:
            out[-1]["recovery"] = recovery_leg(recover_steps)
        log(f"adjoint {name}: {out[-1]}")
    return out


def recovery_leg(steps: int) -> dict:
    """The inverse problem of the differentiable workload: Adam over all 557
    parameters from the nominal set to a 256-spp image of a set with every
    parameter scaled by its own factor in [0.9, 1.1] (PCG seed 2)."""
    from stargaze_amd import diff

    frame = scenes.diff_system(512, 512)
    nominal, target = scenes.diff_system_parameters(frame)
    scene = diff.DiffScene(frame, (512, 512), 64, 4, albedo_bodies=scenes.DIFF_ALBEDO_BODIES,
                           ring_body=scenes.DIFF_RING_BODY, density=True)
    goal = diff.render_converged(scene, target, passes=4)
    t0 = time.perf_counter()
    fitted, losses = diff.recover(scene, goal, nominal, steps=steps, lr=0.01, final_lr=0.001,
                                  log=lambda it, loss: log(f"adam {it}: loss {loss:.4e}"))
    wall = time.perf_counter() - t0
    scene.close()
    before, after = diff.parameter_errors(nominal, target), diff.parameter_errors(fitted, target)
    return {"steps": steps, "parameters": 557, "optimizer": "Adam, relative coordinates, lr 0.01 -> 0.001",
            "target": "256 spp image of the parameters scaled by factors in [0.9, 1.1] (PCG seed 2)",
            "wall_s": round(wall, 2), "ms_per_step": round(wall / steps * 1e3, 2),
            "loss_first": losses[0], "loss_last10": float(np.mean(losses[-10:])),
            "mean_relative_error_before": {k: round(v, 5) for k, v in before.items()},
            "mean_relative_error_after": {k: round(v, 5) for k, v in after.items()}}


# The reference's own suite on an AMD RX 7600 XT (RADV), GPU median ms per
# frame at 1280x720 display, 75% scale, 4 bounces, 70% vegetation, 1 spp per
# submission: benchmarks/optimizations/rx7600xt-2026-10-04/summary.md:12-25.
SUITE_PUBLISHED_RX7600XT = {
    "halo-rings": (34.734, 34.218), "earth-daylight": (120.498, 120.813), "earth-twilight": (41.253, 40.558),
    "median-dense": (74.684, 74.474), "vantus-eclipse": (32.339, 31.490), "vantus-airless": (2.684, 2.549),
    "dual-eclipse": (1.668, 0.960),
}


def suite_leg(repeats: int) -> dict:
    """The reference's render benchmark cases (``render_benchmark``) at its
    published setting: per case and mode the median over ``repeats`` x 16
    timed frames after 8 warm-up submissions. ``gpu_ms`` is the HIP-event
    time of the integrator kernels of one pass; ``wall_ms`` spans the whole
    ``encode`` (exposure metering included) through a stream synchronize."""
    from stargaze_amd import render_benchmark as rb

    cfg = rb.Config(output="", width=1280, height=720, scale=75, bounces=4, vegetation=70, warmup=8, frames=16,
                    repeats=repeats, image_samples=32)
    cases = rb.time_cases(cfg, log=log)
    out = {}
    for case in cases:
        name = case["name"]
        if "skipped" in case:
            out[name] = {"skipped": case["skipped"]}
            continue
        entry = {}
        for mode in case["modes"]:
            runs = [m for r in mode["repeats"] for m in r]
            entry[mode["mode"]] = {"gpu_ms": round(rb.median(m["gpu_ms"] for m in runs), 4),
                                   "wall_ms": round(rb.median(m["wall_ms"] for m in runs), 4)}
        if name in SUITE_PUBLISHED_RX7600XT:
            paused, advancing = SUITE_PUBLISHED_RX7600XT[name]
            entry["rx7600xt_gpu_ms"] = {"paused": paused, "advancing": advancing}
            entry["speedup_wall_vs_rx7600xt_gpu"] = {
                "paused": round(paused / entry["paused"]["wall_ms"], 1),
                "advancing": round(advancing / entry["advancing"]["wall_ms"], 1)}
        out[name] = entry
    return {"setting": "1280x720 display, 960x540 traced, 4 bounces, vegetation 70, 1 spp/submission, "
                       f"warmup 8, {repeats} x 16 frames", "traced_paths_per_frame": 960 * 540,
            "terrain": "the reference's generated terrain per ground (terrain.generate); the alpine ground on a "
                       "seeded stand-in unless STARGAZE_TERRAIN_SURVEYS names the reference's surveys",
            "cases": out}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="halo_4k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default integrator launch")
    ap.add_argument("--launch", default=None,
                    help="integrator launch instead of the autotuner's: fused, or wavefront:CHUNK:TAIL")
    ap.add_argument("--no-work-model", action="store_true", help="skip the oracle event count (no roofline)")
    ap.add_argument("--no-rmse", action="store_true", help="skip the fixed-RMSE leg")
    ap.add_argument("--batch-steps", type=int, default=3, help="passes of the 96-view batch leg (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=3, help="passes of the adjoint leg (0 = skip)")
    ap.add_argument("--recover-steps", type=int, default=200,
                    help="Adam steps of the differentiable workload's inverse problem (0 = skip)")
    ap.add_argument("--suite-repeats", type=int, default=3,
                    help="repeats of the reference's render benchmark cases (0 = skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device buffers 4 times (the reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each": src.numel() * 4, "copies": 4}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build", file=sys.stderr)
        return 2
    builder, width, height, spp, bounces = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    frame, terrain, profile, air = builder(width, height)
    log(f"scene built in {time.perf_counter() - t0:.1f} s")
    tracer = PathTracer((width, height), options=Options(samples_per_frame=spp, max_bounces=bounces, milky_way=0.0))
    if terrain is not None:
        tracer.upload_terrain(*terrain)
    if profile is not None:
        tracer.set_ring_profile(*profile)
    if air is not None:
        tracer.set_air_model(**air)
    log(f"workload {args.workload}: {width}x{height}, {spp} spp/step, {bounces} bounces")
    launch = None
    if args.launch:
        from stargaze_amd.pathtracer import INTEGRATOR_FUSED, INTEGRATOR_WAVEFRONT

        if args.launch == "fused":
            tracer.set_integrator(INTEGRATOR_FUSED)
        else:
            _, chunk, tail = args.launch.split(":")
            tracer.set_integrator(INTEGRATOR_WAVEFRONT, int(chunk))
            tracer.set_tail(int(tail))
    elif not args.no_autotune:
        launch, tuned_ms = tracer.autotune(frame, repeats=1)
        log(f"autotune: {launch} {tuned_ms:.2f} ms")
    for i in range(args.warmup):
        tracer.encode(frame)
    tracer.synchronize()
    log("warmup done")
    tracer.set_profiling(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        tracer.encode(frame)
        log(f"step {i + 1}/{args.steps}")
    tracer.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms, launches = tracer.kernel_time()
    from stargaze_amd.pathtracer import STEP_AIR_OCCLUSION, STEP_CAMERA_AIR

    air_ms, air_launches = tracer.step_time(STEP_CAMERA_AIR)
    occlusion = tracer.step_time(STEP_AIR_OCCLUSION)
    img = tracer.read_accumulation()
    assert np.all(np.isfinite(img)), "non-finite radiance"
    samples_per_step = width * height * spp
    value = samples_per_step * args.steps / elapsed / 1e6
    if args.no_work_model:
        wm = {"flop": float("nan"), "flop_camera_air": float("nan"), "flop_flat_loop_props": float("nan"),
              "flop_camera_air_flat_loop_props": float("nan"), "flop_air_occlusion": float("nan")}
    else:
        log("work model")
        wm = work_model(frame, terrain, profile, air, width, height, spp, bounces)
    kernel_s = kernel_ms / 1e3 / max(launches, 1)
    pass_tflops = wm["flop"] * samples_per_step / kernel_s / 1e12
    compulsory = width * height * 16 * 2 * 2  # accumulation + backdrop read-modify-write
    g = frame.globals
    air_model = air or {"view_steps": 8, "light_steps": 4, "mie_scale_height": 0.0}
    au = scenes.METRES_PER_AU
    config = {"workload": args.workload, "width": width, "height": height, "spp_per_step": spp,
              "max_bounces": bounces}
    if float(g["atmo_params"][3]) >= 0.5:
        config.update({
            "air_view_steps": air_model["view_steps"], "air_light_steps": air_model["light_steps"],
            "rayleigh_scale_height_m": round(float(g["atmo_params"][1]) * au, 1),
            "mie_scale_height_m": round((air_model["mie_scale_height"] or float(g["atmo_params"][1])) * au, 1),
        })
    if terrain is not None:
        kinds = frame.ground["center"][:, 3]
        config.update({"terrain_cells": int(terrain[1]), "rock_boxes": int(np.sum(kinds == 0.0)),
                       "conifer_ellipsoids": int(np.sum(kinds == 1.0)), "ground_records": int(len(frame.ground))})
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpath-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "launch": launch if launch is not None else tracer.launch,
        "roofline": roofline(args.workload, wm, samples_per_step, kernel_s, pass_tflops, compulsory, air_ms,
                             air_launches, tracer.launch, occlusion),
    }
    if args.batch_steps > 0:
        line["batch"] = []
        for name in BATCH_WORKLOADS:
            log(f"batch leg {name}")
            line["batch"].append(batch_leg(args.batch_steps, 1, name))
    if args.adjoint_steps > 0:
        log("adjoint leg")
        line["adjoint"] = adjoint_leg(args.adjoint_steps, args.recover_steps)
    if args.suite_repeats > 0:
        log("reference suite leg")
        line["reference_suite"] = suite_leg(args.suite_repeats)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(frame, terrain, profile, air, width, height, spp, bounces)
    if args.workload == "halo_4k" and not args.no_rmse:
        log("fixed-RMSE leg")
        cpu_rate = line.get("cpu_baseline", {}).get("value")
        line["rmse"] = rmse_leg(elapsed / args.steps * 1e3, spp, width * height, cpu_rate)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
