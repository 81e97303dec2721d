"""Build the HIP library in-tree: ``python -m stargaze_amd.build``.

hipcc cross-compiles for gfx950 without a GPU. The product library is
``stargaze_amd/libstargaze_amd.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
# The host library and the kernel translation units. Each kernel unit
# compiles the integrator for its own kernels, so they build in parallel.
SOURCES = [CSRC / "sg_library.hip", *sorted(CSRC.glob("sg_k_*.hip"))]
SOURCE = SOURCES[0]
OBJECTS = PACKAGE / "_build"
LIBRARY = PACKAGE / "libstargaze_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Keep the compiler from forming packed-FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division and square root to within a few ULP (v_rcp_f32 /
    # v_sqrt_f32 based), the accuracy WGSL promises the reference's shaders
    # (2.5 ULP division), instead of correctly rounded sequences of about
    # ten instructions each: 10% of the 4K Halo pass.
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # No multiply-add contraction beyond explicit fmaf: a step function then
    # rounds the same in every kernel it is inlined into, so the fused,
    # wavefront, tail and batch kernels give bit-identical images (2% of the
    # 4K Halo pass).
    "-ffp-contract=off",
    "-I", str(PACKAGE.parent / "include"),
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}: {cmd[-1]}")


# Forward kernel units built again without the landscape: with the solid-body
# surface styles (suffix _st) and without them (suffix _nl).
PLAIN_VARIANTS = {"sg_k_trace", "sg_k_batch", "sg_k_wave_a", "sg_k_wave_b"}


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every translation unit (in parallel) and link the library."""
    from concurrent.futures import ThreadPoolExecutor

    OBJECTS.mkdir(exist_ok=True)
    units = [(src, OBJECTS / (src.stem + ".o"), []) for src in SOURCES]
    units += [(src, OBJECTS / (src.stem + "_nl.o"), ["-DSG_LANDSCAPE=0", "-DSG_STYLES=0"])
              for src in SOURCES if src.stem in PLAIN_VARIANTS]
    units += [(src, OBJECTS / (src.stem + "_st.o"), ["-DSG_LANDSCAPE=0"])
              for src in SOURCES if src.stem in PLAIN_VARIANTS]
    # Longest first: the adjoint and batch units dominate.
    units.sort(key=lambda u: ("adjoint" not in u[0].stem, "batch" not in u[0].stem, "trace" not in u[0].stem))
    objects = [u[1] for u in units]
    jobs = jobs or min(len(units), max(1, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(lambda u: _run([HIPCC, *FLAGS, *u[2], "-c", "-o", str(u[1]), str(u[0])], verbose), units))
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)], verbose)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
