// Exposure, display, terrain and sky caches, the probe kernel, and the reset and
// advance nodes of a captured batch pass.
#include "sg_launch.h"

namespace sgk {

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Luminance histogram of the central rectangle (exposure.wgsl:51-76): each
// workgroup bins its 16x16 tile in LDS and adds its non-empty bins to the
// device histogram.
__global__ __launch_bounds__(256) void k_meter(MeterArgs ma) {
    __shared__ uint32_t s_bins[kHistogramBins];
    const uint32_t t = threadIdx.y * 16u + threadIdx.x;
    s_bins[t] = 0u;
    __syncthreads();
    const uint32_t x = blockIdx.x * 16u + threadIdx.x, y = blockIdx.y * 16u + threadIdx.y;
    if (x < ma.meter_w && y < ma.meter_h) {
        const float4 hdr = ma.accumulation[(size_t)(y + ma.origin_y) * ma.width + (x + ma.origin_x)];
        // Unfused products and sums, as written (the oracle's rounding).
        const float dot = __fadd_rn(__fadd_rn(__fmul_rn(hdr.x, 0.2126f), __fmul_rn(hdr.y, 0.7152f)),
                                    __fmul_rn(hdr.z, 0.0722f));
        const float luminance = fmaxf(0.0f, __fsub_rn(dot, hdr.w));
        float safe = luminance;
        if (!(safe > 0.0f)) safe = exp2f(kLogMin);
        safe = fminf(fmaxf(safe, exp2f(kLogMin)), exp2f(kLogMax));
        const float u = (log2f(safe) - kLogMin) / (kLogMax - kLogMin);
        const uint32_t bin = (uint32_t)fminf(fmaxf(u * (float)kHistogramBins, 0.0f), (float)(kHistogramBins - 1u));
        atomicAdd(&s_bins[bin], 1u);
    }
    __syncthreads();
    if (s_bins[t] != 0u) atomicAdd(&ma.histogram[t], s_bins[t]);
}

// Percentile of the histogram and the eased automatic exposure
// (exposure.wgsl:78-109), one workgroup of kHistogramBins lanes: an
// inclusive scan in LDS replaces the serial cumulative loop.
__global__ __launch_bounds__(256) void k_resolve(const uint32_t *histogram, sg_exposure *state) {
    __shared__ uint32_t s_scan[kHistogramBins];
    __shared__ uint32_t s_bin;
    const uint32_t t = threadIdx.x;
    const float enabled = state->enabled;
    s_scan[t] = histogram[t];
    if (t == 0) s_bin = kHistogramBins;
    __syncthreads();
    for (uint32_t offset = 1; offset < kHistogramBins; offset <<= 1) {
        const uint32_t add = t >= offset ? s_scan[t - offset] : 0u;
        __syncthreads();
        s_scan[t] += add;
        __syncthreads();
    }
    const uint32_t total = s_scan[kHistogramBins - 1u];
    if (enabled < 0.5f || total == 0u) return;
    const float percentile = fminf(fmaxf(state->percentile, 0.0f), 1.0f);
    const uint32_t cutoff = umax(1u, (uint32_t)ceilf((float)total * percentile));
    const uint32_t before = t > 0 ? s_scan[t - 1u] : 0u;
    if (s_scan[t] >= cutoff && before < cutoff) s_bin = t;
    __syncthreads();
    if (t != 0) return;
    const uint32_t bin = s_bin < kHistogramBins ? s_bin : 0u;
    const float u = ((float)bin + 0.5f) / (float)kHistogramBins;
    const float measured = fmaxf(exp2f(kLogMin + u * (kLogMax - kLogMin)), kMeterFloor);
    const float desired = fminf(fmaxf(state->key / measured, kMinExposure), kMaxExposure);
    const float rate = 0.15f;
    state->value = fminf(fmaxf(state->value + (desired - state->value) * rate, kMinExposure), kMaxExposure);
}

__device__ __forceinline__ float4 hdr_at(const float4 *img, int x, int y, int w, int h) {
    x = imin(imax(x, 0), w - 1);
    y = imin(imax(y, 0), h - 1);
    return img[(size_t)y * w + x];
}
__device__ __forceinline__ float aces(float x) {
    return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f), 1.0f);
}
// IEC 61966-2-1 sRGB transfer function, then 8-bit rounding (what an
// Rgba8UnormSrgb target stores for a fragment value v).
__device__ __forceinline__ uint8_t srgb8(float v) {
    const float e = v <= 0.0031308f ? 12.92f * v : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
    return (uint8_t)fminf(fmaxf(e * 255.0f + 0.5f, 0.0f), 255.0f);
}

// display.wgsl:62-97 for every pixel of an out_w x out_h target.
__global__ __launch_bounds__(256) void k_display(DisplayArgs da) {
    const uint32_t x = blockIdx.x * 16u + threadIdx.x, y = blockIdx.y * 16u + threadIdx.y;
    if (x >= da.out_w || y >= da.out_h) return;
    const int w = (int)da.width, h = (int)da.height;
    const float px = ((float)x + 0.5f) / (float)da.out_w * (float)da.width - 0.5f;
    const float py = ((float)y + 0.5f) / (float)da.out_h * (float)da.height - 0.5f;
    const int lx = (int)floorf(px), ly = (int)floorf(py);
    const float fx = px - floorf(px), fy = py - floorf(py);
    auto lerp4 = [](float4 a, float4 b, float t) {
        return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
    };
    const float4 hdr = lerp4(lerp4(hdr_at(da.accumulation, lx, ly, w, h), hdr_at(da.accumulation, lx + 1, ly, w, h), fx),
                             lerp4(hdr_at(da.accumulation, lx, ly + 1, w, h), hdr_at(da.accumulation, lx + 1, ly + 1, w, h), fx),
                             fy);
    const float4 sky = lerp4(lerp4(hdr_at(da.backdrop, lx, ly, w, h), hdr_at(da.backdrop, lx + 1, ly, w, h), fx),
                             lerp4(hdr_at(da.backdrop, lx, ly + 1, w, h), hdr_at(da.backdrop, lx + 1, ly + 1, w, h), fx),
                             fy);
    const sg_exposure st = *da.state;
    const bool automatic = st.enabled > 0.5f;
    const float base = automatic ? st.value : da.manual_exposure;
    const float stops = fminf(fmaxf(log2f(fminf(fmaxf(base, 1.0e-30f), 1.0e30f)) + fminf(fmaxf(st.bias, -100.0f), 100.0f),
                                    -100.0f), 100.0f);
    const float multiplier = exp2f(stops);
    const float cap = 1.0e6f / multiplier;
    const float backdrop_gain = fminf(multiplier, 0.04f / fmaxf(da.milky_way, 1.0e-30f));
    const float r = fminf(fmaxf(hdr.x - sky.x, 0.0f), cap) * multiplier + sky.x * backdrop_gain;
    const float g = fminf(fmaxf(hdr.y - sky.y, 0.0f), cap) * multiplier + sky.y * backdrop_gain;
    const float b = fminf(fmaxf(hdr.z - sky.z, 0.0f), cap) * multiplier + sky.z * backdrop_gain;
    const float4 mapped = make_float4(aces(r), aces(g), aces(b), 1.0f);
    const size_t i = (size_t)y * da.out_w + x;
    if (da.out_linear) da.out_linear[i] = mapped;
    if (da.out_srgb) da.out_srgb[i] = make_uchar4(srgb8(mapped.x), srgb8(mapped.y), srgb8(mapped.z), 255);
}

// cache_terrain_geometry (src/shaders/terrain.wgsl:242-285), one lane per index.
__global__ void k_terrain_cache(Scene sc, float *terrain) {
    terrain_cache_entry(sc, terrain, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// cache_surface_sky (src/shaders/forest.wgsl:3-17): 16x8 texels.
__global__ void k_sky_cache(Scene sc, float4 *sky) {
    const uint32_t index = threadIdx.x;
    if (index >= 128u) return;
    f3 light = sky_cache_texel(sc, index % 16u, index / 16u);
    sky[index] = make_float4(light.x, light.y, light.z, 0.f);
}

// Probe kernel: production helpers on caller-supplied rays.
__global__ void k_probe(Scene sc, uint32_t kind, const sg_probe_ray *rays, uint32_t count, sg_probe_result *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const sg_probe_ray pr = rays[i];
    Ray ray{d3{pr.offset[0], pr.offset[1], pr.offset[2]}, f3{pr.direction[0], pr.direction[1], pr.direction[2]},
            pr.anchor};
    sg_probe_result r;
    memset(&r, 0, sizeof(r));
    r.index = SG_MISS;
    r.distance = 1.0e30f;
    switch (kind) {
    case SG_PROBE_SPHERE:
    case SG_PROBE_CLOSEST:
    case SG_PROBE_LANDSCAPE: {
        Hit h = kind == SG_PROBE_SPHERE ? sphere_hit(sc, ray, pr.target)
                                        : (kind == SG_PROBE_CLOSEST ? closest_hit(sc, ray) : landscape_hit(sc, ray));
        r.distance = h.distance;
        r.index = h.index;
        r.normal[0] = h.normal.x; r.normal[1] = h.normal.y; r.normal[2] = h.normal.z;
        break;
    }
    case SG_PROBE_RING: {
        RingHit h = ring_hit(sc, ray, pr.target);
        r.distance = h.distance;
        r.index = h.index;
        r.extra[0] = h.point.x; r.extra[1] = h.point.y; r.extra[2] = h.point.z;
        if (h.index != SG_MISS) {
            r.normal[0] = ring_transmission(sc, ray, h);
            r.normal[1] = ring_optics(sc, h.index, h.point).w;
        }
        break;
    }
    case SG_PROBE_VISIBILITY:
        r.distance = light_visibility(sc, ray, pr.target);
        r.index = 0;
        break;
    case SG_PROBE_TERRAIN: {
        f3 o = tof(ray.offset);
        float4 h = terrain_query(sc, o, ray.direction, 1.0e30f, pr.target != 0u);
        r.distance = h.w;
        r.index = h.w < 1.0e29f ? SG_TERRAIN_ID : SG_MISS;
        r.normal[0] = h.x; r.normal[1] = h.y; r.normal[2] = h.z;
        break;
    }
    case SG_PROBE_ALBEDO: {
        const Body &b = body(sc, pr.target);
        const f3 a = surface_style(b) > 0u ? style_albedo(b, ray.direction) : surface_albedo(b, ray.direction);
        r.index = 0u;
        r.normal[0] = a.x; r.normal[1] = a.y; r.normal[2] = a.z;
        break;
    }
    case SG_PROBE_CONE: {
        r.distance = cone_width(sc, ray, pr.anchor == SG_MISS ? 0u : pr.anchor);
        break;
    }
    case SG_PROBE_NOISE: {
        r.index = 0u;
        r.distance = vnoise(ray.direction);
        r.normal[0] = hash31(ray.direction);
        r.normal[1] = fbm(ray.direction, pr.target);
        break;
    }
    case SG_PROBE_SKY: {
        uint32_t rng = hash_u32(pr.target + 817u);
        const f3 light = surface_sky(sc, ray, ray.direction, rng);
        const f3 texel = surface_sky_texel(sc, (int)(pr.target % 16u), (int)((pr.target / 16u) % 8u));
        r.index = 0u;
        r.normal[0] = light.x; r.normal[1] = light.y; r.normal[2] = light.z;
        r.extra[0] = texel.x; r.extra[1] = texel.y; r.extra[2] = texel.z;
        break;
    }
    case SG_PROBE_GROUND: {
        const float distance = __uint_as_float(pr.target);
        const f3 p = {pr.offset[0], pr.offset[1], pr.offset[2]};
        const f3 shading = ground_shading_normal(sc, p, distance, from_ground(sc, ray.direction), pr.anchor);
        const f3 a = ground_albedo(sc, p, distance, ray.direction, pr.anchor);
        r.index = pr.anchor;
        r.distance = distance;
        r.normal[0] = shading.x; r.normal[1] = shading.y; r.normal[2] = shading.z;
        r.extra[0] = a.x; r.extra[1] = a.y; r.extra[2] = a.z;
        break;
    }
    default:
        break;
    }
    out[i] = r;
}

__global__ void k_batch_reset(BatchArgs ba) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ba.counter[0] = 0u;
        ba.clock[0] = ~0ull;
        ba.clock[1] = 0ull;
    }
    if (i < ba.entries) {
        ba.done[i] = 0ull;
    }
}

// End of a pass: each entry's accumulated count and noise sequence advance
// as the host history does, so an unchanged table can be replayed.
__global__ void k_batch_advance(BatchArgs ba) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ba.entries) return;
    BatchView &v = ba.table[e];
    v.ta.accumulated += v.ta.count;
    v.ta.sequence += v.ta.count;
    v.scene.s.counts[1] += v.ta.count;
    v.scene.s.sampling[0] += v.ta.count;
}

}  // namespace sgk
