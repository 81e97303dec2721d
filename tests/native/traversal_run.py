"""Runs the host build of the product's landscape and terrain traversals on
the Halo parity scene and writes their answers; run in a subprocess so a
traversal that fails to terminate ends in a timeout, not a stuck suite."""
import ctypes
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "oracle"), str(ROOT / "tests" / "native")]
import build  # noqa: E402
import oracle  # noqa: E402
from stargaze_amd import scenes  # noqa: E402
from stargaze_amd.layouts import MISS, PROBE_RAY, PROBE_RESULT  # noqa: E402
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


def main(out_path):
    lib = ctypes.CDLL(str(build.build()))
    frame, data, cells = scenes.halo_parity(32, 18)
    orc = oracle.OracleTracer((4, 4), Options(milky_way=0.0), "f32")
    orc.upload_terrain(data, cells)
    rng = np.random.default_rng(11)
    n = 4096
    rays = np.zeros(n, PROBE_RAY)
    dirs = rng.normal(size=(n, 3))
    dirs[: n // 2, 1] = -np.abs(dirs[: n // 2, 1]) * 0.3 - 0.02
    g = frame.globals
    east, up, north = (g[k][:3].astype(np.float64) for k in ("ground_east", "ground_up", "ground_north"))
    world = dirs[:, :1] * east + dirs[:, 1:2] * up + dirs[:, 2:3] * north
    world /= np.linalg.norm(world, axis=1, keepdims=True)
    rays["anchor"] = MISS
    rays["direction"] = world
    rays["target"][n // 4: n // 2] = 1  # occlusion queries
    ref = orc.probe(frame, 5, rays)  # also fills the oracle's terrain cache
    st = orc.settings(frame, 0)
    cap = 4 * len(frame.ground) + 8
    nodes = np.zeros((cap, 8), np.float32)
    prims = np.zeros(cap, np.uint32)
    pc = ctypes.c_uint32()
    ground = np.ascontiguousarray(frame.ground)
    lib.hc_build_bvh.restype = ctypes.c_uint32
    count = lib.hc_build_bvh(ground.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(len(ground)),
                             nodes.ctypes.data_as(ctypes.c_void_p), prims.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_uint32(cap), ctypes.byref(pc))
    assert count <= cap
    bodies = np.ascontiguousarray(frame.bodies)
    terrain = product_terrain(lib, st, data)
    results = {}
    for name, use_bvh in (("bvh", True), ("flat", False)):
        out = np.zeros(n, PROBE_RESULT)
        lib.hc_query(st.ctypes.data_as(ctypes.c_void_p), bodies.ctypes.data_as(ctypes.c_void_p),
                     ground.ctypes.data_as(ctypes.c_void_p), terrain.ctypes.data_as(ctypes.c_void_p),
                     nodes.ctypes.data_as(ctypes.c_void_p) if use_bvh else None,
                     prims.ctypes.data_as(ctypes.c_void_p) if use_bvh else None,
                     ctypes.c_uint32(5), rays.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(n),
                     out.ctypes.data_as(ctypes.c_void_p))
        results[name] = out
    np.savez(out_path, bvh=results["bvh"], flat=results["flat"], ref_index=ref["index"],
             ref_distance=ref["distance"], occlusion=rays["target"])


if __name__ == "__main__":
    main(sys.argv[1])
