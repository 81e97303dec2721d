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
    if constexpr (kLandscape) {
        // Shadow rays with landscape predicates owed go to k_air_occlusion;
        // k_air_resolve adds what they carry, scaled by this throughput.
        const f3 scale = st.throughput;
        AirDefer defer{wa.air_a, wa.air_b, wa.air_c, wa.air_count, wa.air_capacity, SG_MISS};
        path_camera_air<NoGrad, AirDefer>(sc, st, h, NoGrad(), &defer);
        wa.air_pending[id] = make_float4(scale.x, scale.y, scale.z, __uint_as_float(defer.head));
    } else {
        path_camera_air(sc, st, h);
    }
    wa.thr[id] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, t.w);
    wa.rad[id] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, r.w);
    wa.rbase[id] = make_float4(b.x, b.y, b.z, __uint_as_float((__float_as_uint(b.w) & ~255u) | st.flags));
    wa.rng[id] = st.rng;
}

// The landscape predicates the camera-air kernel deferred, one lane per
// queued shadow ray (blockIdx.y: queue region; grid-stride over its count):
// a blocked ray's contribution is cleared.
__global__ SG_WAVES(SG_EXTEND_WAVES) void SG_K(k_air_occlusion)(Scene sc, WaveArgs wa) {
    const uint32_t n = umin(wa.air_count[blockIdx.y], wa.air_capacity);
    const uint32_t base = blockIdx.y * wa.air_capacity;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = base + i;
        const float4 a = wa.air_a[r], b = wa.air_b[r];
        const Ray ray{d3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, SG_MISS};
        if (air_light_blocked(sc, ray, a.w, __float_as_uint(wa.air_c[r].w)))
            wa.air_c[r] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Each path adds the contributions of its deferred shadow rays (latest
// first), scaled by its throughput at the camera-air step.
__global__ __launch_bounds__(256) void SG_K(k_air_resolve)(WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    const float4 pending = wa.air_pending[id];
    uint32_t r = __float_as_uint(pending.w);
    if (r == SG_MISS) return;
    f3 sum = f3s(0.0f);
    while (r != SG_MISS) {
        const float4 c = wa.air_c[r];
        sum += f3{c.x, c.y, c.z};
        r = __float_as_uint(wa.air_b[r].w);
    }
    float4 rad = wa.rad[id];
    const f3 add = f3{pending.x, pending.y, pending.z} * sum;
    wa.rad[id] = make_float4(rad.x + add.x, rad.y + add.y, rad.z + add.z, rad.w);
}

}  // namespace sgk
