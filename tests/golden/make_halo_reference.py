"""Writes the Halo RMSE reference: ``halo_rmse_reference.npz``.

The 4K Halo workload's camera, scene and air model (16 x 8 march, 1.2 km Mie
scale height) at 320 x 180 pixels, rendered on the GPU with 16384 samples per
pixel: 256 fresh passes of 64 samples (``PathTracer.restart`` before each, so
every pass draws new samples), their means summed in float64 on the host and
stored as float32. Noise sequences start at ``REFERENCE_SEED``, far from the
sequences the benchmark's own renders use, so the two share no draws.

Run on a GPU: ``python tests/golden/make_halo_reference.py``. The target RMSE
(the oracle's at 32 spp) is added by ``--target`` on any host.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

WIDTH, HEIGHT = 320, 180
PASSES, PASS_SPP = 256, 64
REFERENCE_SEED = 0x40000000
TARGET_SPP = 32
OUT = HERE / "halo_rmse_reference.npz"
META = HERE / "halo_rmse_reference.json"



def scene():
    from stargaze_amd import scenes

    return scenes.halo_observatory(WIDTH, HEIGHT)


def render_reference(out_dir: Path = HERE) -> None:
    from stargaze_amd import scenes
    from stargaze_amd.pathtracer import Options, PathTracer

    frame, data, cells = scene()
    t = PathTracer((WIDTH, HEIGHT), options=Options(noise_seed=REFERENCE_SEED, samples_per_frame=PASS_SPP,
                                                    max_bounces=4, milky_way=0.0))
    t.upload_terrain(data, cells)
    t.set_air_model(**scenes.HALO_AIR)
    total = np.zeros((HEIGHT, WIDTH, 3), np.float64)
    t0 = time.perf_counter()
    for i in range(PASSES):
        t.restart()
        t.encode(frame)
        total += t.read_accumulation()[..., :3].astype(np.float64)
        if i % 32 == 31:
            print(f"pass {i + 1}/{PASSES} {time.perf_counter() - t0:.1f} s", flush=True)
    t.close()
    image = (total / PASSES).astype(np.float32)
    out_dir.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(out_dir / OUT.name, image=image)
    (out_dir / META.name).write_text(json.dumps({
        "scene": "halo_observatory (4096-cell terrain, 20k boxes, 50k ellipsoids) with HALO_AIR",
        "width": WIDTH, "height": HEIGHT, "spp": PASSES * PASS_SPP, "max_bounces": 4,
        "noise_seed": REFERENCE_SEED, "accumulation": "float64 sum of 256 fresh 64-spp pass means",
    }, indent=1) + "\n")
    print(f"wrote {out_dir / OUT.name}")


# Noise seeds far apart: a pass of n samples draws sequences seed .. seed + n - 1,
# so nearby seeds would share most of their draws.
TARGET_SEEDS = tuple(k * 1_000_003 for k in range(8))


def oracle_target() -> None:
    """The FP32 oracle's expected linear-RGB RMSE against the reference at
    TARGET_SPP spp: the root mean square of its RMSE over eight independent
    noise seeds.
    This is the fixed RMSE the benchmark's time-to-RMSE converges to."""
    import oracle
    from stargaze_amd import scenes
    from stargaze_amd.pathtracer import Options

    sys.path.insert(0, str(ROOT / "tests"))
    import parity

    frame, data, cells = scene()
    ref = np.load(OUT)["image"]
    per_seed = []
    t0 = time.perf_counter()
    for seed in TARGET_SEEDS:
        t = oracle.OracleTracer((WIDTH, HEIGHT), Options(noise_seed=seed, samples_per_frame=TARGET_SPP,
                                                         max_bounces=4, milky_way=0.0))
        t.upload_terrain(data, cells)
        t.set_air_model(**scenes.HALO_AIR)
        t.encode(frame)
        per_seed.append(parity.rmse(t.read_accumulation(), ref))
    dt = time.perf_counter() - t0
    target = float(np.sqrt(np.mean(np.square(per_seed))))
    meta = json.loads(META.read_text())
    meta["target"] = {"rmse": target, "spp": TARGET_SPP, "noise_seeds": list(TARGET_SEEDS),
                      "rmse_per_seed": per_seed,
                      "source": "FP32 oracle, linear RGB (W excluded), root mean square over the seeds"}
    META.write_text(json.dumps(meta, indent=1) + "\n")
    print(f"oracle RMSE at {TARGET_SPP} spp: {target:.6g} (seeds {per_seed}, {dt:.1f} s)")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--target", action="store_true", help="add the oracle's target RMSE to the metadata")
    ap.add_argument("--out-dir", type=Path, default=HERE, help="where the rendered reference goes")
    args = ap.parse_args()
    oracle_target() if args.target else render_reference(args.out_dir)
