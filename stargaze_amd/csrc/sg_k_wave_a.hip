// Wavefront integrator kernels, first half: generate, extend, camera air.
#include "sg_wave.h"

namespace sgk {

__global__ __launch_bounds__(256) void SG_K(k_generate)(Scene sc, WaveArgs wa, uint32_t *queue) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= wa.paths) return;
    const uint32_t pixel = wa.pixel_begin + id / wa.count;
    const uint32_t s = id % wa.count;
    const uint32_t x = pixel % wa.width, y = pixel / wa.width;
    uint32_t rng = path_seed(pixel, wa.sequence + s);
    f3 direction = camera_direction(sc, x, y, rng);
    PathState st;
    path_begin(st, direction, rng, path_accumulated(wa, id));
    store_state(wa, id, st);
    queue[id] = id;
}

__global__ SG_WAVES(SG_EXTEND_WAVES) void SG_K(k_extend)(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    PathState st;
    load_ray(wa.ray_o[id], wa.ray_d[id], st.ray);
    uint32_t packed = __float_as_uint(wa.rbase[id].w);
    st.depth = (packed >> 8) & 255u;
    st.null_crossings = packed >> 16;
    st.rng = wa.rng[id];
    HitRecord h = path_extend(sc, st);
    wa.rng[id] = st.rng;
    wa.h0[id] = make_float4(h.hit.distance, __uint_as_float(h.hit.index), h.disc_scale, h.ring.distance);
    wa.h1[id] = make_float4(h.hit.normal.x, h.hit.normal.y, h.hit.normal.z, __uint_as_float(h.ring.index));
    wa.h2[id] = make_float4(h.ring.point.x, h.ring.point.y, h.ring.point.z, 0.f);
}

__global__ SG_WAVES(SG_AIR_WAVES) void SG_K(k_camera_air)(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    PathState st;
    float4 d = wa.ray_d[id];
    load_ray(wa.ray_o[id], d, st.ray);
    float4 t = wa.thr[id], r = wa.rad[id], b = wa.rbase[id];
    st.throughput = f3{t.x, t.y, t.z};
    st.radiance = f3{r.x, r.y, r.z};
    st.flags = __float_as_uint(b.w) & 255u;
    st.rng = wa.rng[id];
    st.accumulated = path_accumulated(wa, id);
    HitRecord h;
    load_hit(wa, id, h);
    path_camera_air(sc, st, h);
    wa.thr[id] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, t.w);
    wa.rad[id] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, r.w);
    wa.rbase[id] = make_float4(b.x, b.y, b.z, __uint_as_float((__float_as_uint(b.w) & ~255u) | st.flags));
    wa.rng[id] = st.rng;
}

}  // namespace sgk
