// Structure-of-arrays path state of the wavefront integrator: loads and
// stores shared by its kernels (sg_k_wave_a.hip, sg_k_wave_b.hip).
#pragma once
#include "sg_launch.h"

namespace sgk {

SG_FN double bits_d(uint32_t u) { return (double)u; }
SG_FN uint32_t d_bits(double d) { return (uint32_t)d; }

static __device__ __forceinline__ uint32_t path_accumulated(const WaveArgs &wa, uint32_t id) {
    return wa.has_air ? wa.accumulated + id % wa.count : wa.accumulated;
}
static __device__ __forceinline__ void load_ray(const double4 &o, const float4 &d, Ray &r) {
    r.offset = d3{o.x, o.y, o.z};
    r.anchor = d_bits(o.w);
    r.direction = f3{d.x, d.y, d.z};
}
static __device__ __forceinline__ void load_state(const WaveArgs &wa, uint32_t id, PathState &st) {
    float4 d = wa.ray_d[id];
    load_ray(wa.ray_o[id], d, st.ray);
    st.previous_pdf = d.w;
    double4 po = wa.prev_o[id];
    st.previous_vertex.offset = d3{po.x, po.y, po.z};
    st.previous_vertex.anchor = d_bits(po.w);
    st.previous_vertex.direction = st.ray.direction;
    float4 t = wa.thr[id];
    st.throughput = f3{t.x, t.y, t.z};
    st.ring_cap = t.w;
    float4 r = wa.rad[id];
    st.radiance = f3{r.x, r.y, r.z};
    st.shine_body = __float_as_uint(r.w);
    float4 b = wa.rbase[id];
    st.ring_base = f3{b.x, b.y, b.z};
    uint32_t packed = __float_as_uint(b.w);
    st.flags = packed & 255u;
    st.depth = (packed >> 8) & 255u;
    st.null_crossings = packed >> 16;
    st.rng = wa.rng[id];
    st.backdrop = f3s(0.0f);
    st.accumulated = path_accumulated(wa, id);
}
static __device__ __forceinline__ void store_state(const WaveArgs &wa, uint32_t id, const PathState &st) {
    wa.ray_o[id] = make_double4(st.ray.offset.x, st.ray.offset.y, st.ray.offset.z, bits_d(st.ray.anchor));
    wa.ray_d[id] = make_float4(st.ray.direction.x, st.ray.direction.y, st.ray.direction.z, st.previous_pdf);
    wa.prev_o[id] = make_double4(st.previous_vertex.offset.x, st.previous_vertex.offset.y,
                                 st.previous_vertex.offset.z, bits_d(st.previous_vertex.anchor));
    wa.thr[id] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, st.ring_cap);
    wa.rad[id] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, __uint_as_float(st.shine_body));
    uint32_t packed = (st.flags & 255u) | ((st.depth & 255u) << 8) | (st.null_crossings << 16);
    wa.rbase[id] = make_float4(st.ring_base.x, st.ring_base.y, st.ring_base.z, __uint_as_float(packed));
    wa.rng[id] = st.rng;
}
static __device__ __forceinline__ void load_hit(const WaveArgs &wa, uint32_t id, HitRecord &h) {
    float4 a = wa.h0[id], b = wa.h1[id], c = wa.h2[id];
    h.hit.distance = a.x;
    h.hit.index = __float_as_uint(a.y);
    h.disc_scale = a.z;
    h.ring.distance = a.w;
    h.hit.normal = f3{b.x, b.y, b.z};
    h.ring.index = __float_as_uint(b.w);
    h.ring.point = f3{c.x, c.y, c.z};
}
// The deferral sink of the camera-air kernel: queues shadow rays whose
// landscape predicates are owed, one atomic per wavefront (ballot and
// prefix of the lanes that push at this call) on its region's counter, and
// links each path's records through `head` (latest first). The counters
// count every push; pushes at or past `capacity` in a region are dropped,
// and the host, seeing a count above `capacity`, grows the queue and runs
// the chunk again from its first kernel.
struct AirDefer {
    static constexpr bool on = true;
    float4 *a, *b, *c;
    uint32_t *count;
    uint32_t capacity;
    uint32_t head;
    __device__ void push(f3 origin, f3 direction, float limit, f3 term, uint32_t owed) {
        const uint64_t mask = __ballot(1);
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t base = 0;
        const uint32_t region = blockIdx.x % kAirRegions;
        if (lane == leader) base = atomicAdd(count + region, (uint32_t)__popcll(mask));
        base = __shfl(base, (int)leader);
        const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (slot >= capacity) return;
        const uint32_t r = region * capacity + slot;
        a[r] = make_float4(origin.x, origin.y, origin.z, limit);
        b[r] = make_float4(direction.x, direction.y, direction.z, __uint_as_float(head));
        c[r] = make_float4(term.x, term.y, term.z, __uint_as_float(owed));
        head = r;
    }
};

static __device__ __forceinline__ void finish(const WaveArgs &wa, uint32_t id, const PathState &st) {
    PathResult r = path_finish(st);
    wa.value[id] = r.value;
    wa.backdrop[id] = make_float4(r.backdrop.x, r.backdrop.y, r.backdrop.z, 0.f);
    if (wa.records)
        wa.records[(size_t)wa.pixel_begin * wa.count + id] =
            make_float4(r.value.x, r.value.y, r.value.z, ring_cap_scale(st));
}

}  // namespace sgk
