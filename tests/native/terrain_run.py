"""Host build of the product's terrain traversal on the reference's terrain
oracle rays; run in a subprocess so a non-terminating traversal times out."""
import ctypes
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "oracle"), str(ROOT / "tests"), str(ROOT / "tests" / "native")]
import build  # noqa: E402
import oracle  # noqa: E402
import terrain_cases as tc  # noqa: E402
from stargaze_amd.layouts import PROBE_RESULT  # noqa: E402
from stargaze_amd.pathtracer import Options  # noqa: E402


def product_terrain(lib, st, data):
    """Terrain buffer in the product's layout, cache filled by the product's
    cache pass run on the host."""
    d, t = ctypes.c_uint64(), ctypes.c_uint64()
    lib.hc_terrain_layout(st.ctypes.data_as(ctypes.c_void_p), ctypes.byref(d), ctypes.byref(t))
    assert d.value == data.size
    buf = np.zeros(t.value, np.float32)
    buf[: data.size] = data
    lib.hc_terrain_cache(st.ctypes.data_as(ctypes.c_void_p), buf.ctypes.data_as(ctypes.c_void_p))
    return buf


def main(out_path, cells, radius):
    lib = ctypes.CDLL(str(build.build()))
    heights, data = tc.make_terrain(cells)
    rays = tc.probe_rays(radius, extra=300)
    frame = tc.probe_frame(cells, radius)
    orc = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f32")
    orc.upload_terrain(data, cells)
    orc.probe(frame, 4, rays[:1])  # fills the terrain cache
    st = orc.settings(frame, 0)
    terrain = product_terrain(lib, st, data)
    out = np.zeros(rays.size, PROBE_RESULT)
    body = np.zeros(160, np.uint8)
    lib.hc_query(st.ctypes.data_as(ctypes.c_void_p), body.ctypes.data_as(ctypes.c_void_p), None,
                 terrain.ctypes.data_as(ctypes.c_void_p), None, None, ctypes.c_uint32(4),
                 rays.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(rays.size), out.ctypes.data_as(ctypes.c_void_p))
    np.savez(out_path, result=out, expected=tc.exhaustive(heights, cells, radius, rays))


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]), float(sys.argv[3]))
