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
SOURCE = PACKAGE / "csrc" / "sg_kernels.hip"
LIBRARY = PACKAGE / "libstargaze_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Keep the compiler from forming packed-FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(PACKAGE.parent / "include"),
]


def build(verbose: bool = False) -> Path:
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), str(SOURCE)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
