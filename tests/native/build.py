"""Host build of the product's device helpers for CPU tests (hipcc -x c++)."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "libhostcheck.so"


def build() -> Path:
    import os

    if os.environ.get("HOSTCHECK_LIB"):
        return Path(os.environ["HOSTCHECK_LIB"])
    OUT.parent.mkdir(exist_ok=True)
    src = HERE / "hostcheck.cpp"
    deps = [src, *(HERE.parent.parent / "stargaze_amd" / "csrc").glob("*.h")]
    if OUT.exists() and all(OUT.stat().st_mtime >= d.stat().st_mtime for d in deps):
        return OUT
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-O2",
                    "-std=c++17", "-fPIC", "-shared", "-I", str(HERE.parent.parent / "include"), "-o", str(OUT),
                    str(src)], check=True, capture_output=True)
    return OUT
