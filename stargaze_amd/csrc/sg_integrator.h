// Atmosphere, sky cache, landscape materials and the path integrator
// (src/shaders/pathtrace.wgsl:927-1626, forest.wgsl:1-54, ground.wgsl:256-666).
#pragma once
#include "sg_geometry.h"

namespace sg {

// --- gradient accumulation for the adjoint ----------------------------------
// Slots of the gradient vector (the C ABI's SG_GRAD_* layout): material rgb
// per body, one per ring profile bin, Rayleigh rgb, Mie and Mie g, then one
// per density profile bin.
constexpr uint32_t GRAD_ALBEDO = 0u;
constexpr uint32_t GRAD_RING = 3u * SG_BODY_CAPACITY;
constexpr uint32_t GRAD_AIR = GRAD_RING + SG_RING_PROFILE_BINS;
constexpr uint32_t GRAD_DENSITY = GRAD_AIR + 5u;
constexpr uint32_t GRAD_COUNT = GRAD_DENSITY + SG_DENSITY_BINS;
constexpr uint32_t GRAD_PRIVATE = SG_DENSITY_BINS + 5u;  // lane-private slots per path

// Profile bin of an altitude: SG_DENSITY_BINS equal slices up to the top of
// the air.
SG_FN uint32_t density_bin(const Scene &sc, float altitude) {
    return umin((uint32_t)fmaxf(altitude / sc.s.globals.atmo_params[2] * (float)SG_DENSITY_BINS, 0.0f),
                SG_DENSITY_BINS - 1u);
}
// Where the samples of a density column fell (for the adjoint): the profile
// bin of each sample, four bits apiece (at most 16 samples), their spacing
// and their sum.
struct ColumnBins {
    float total = 0.0f;
    float dt = 0.0f;
    uint64_t bins = 0u;
    uint32_t count = 0u;
};
constexpr uint32_t COLUMN_BIN_SLOTS = 16u;

// The plain integrator: every gradient hook compiles away.
struct NoGrad {
    static constexpr bool on = false;
    static constexpr bool capture = false;
};
// A primal run for the adjoint: no derivatives, but the ring indirect cap's
// make-up is kept where the cap is set (path_shade): the event (depth and
// null crossings), which term bound it, and the colour of that term.
struct CapTrace {
    static constexpr bool on = false;
    static constexpr bool capture = true;
    mutable bool set = false;
    mutable bool via_direct = false;
    mutable uint32_t depth = 0u, nulls = 0u;
    mutable f3 vector = f3{0.0f, 0.0f, 0.0f};
};
// Camera-air shadow rays answer their landscape queries inline (NoDefer), or
// hand them to a deferral sink that queues them for a kernel of its own (the
// wavefront integrator's AirDefer): push(origin, direction, limit, term,
// owed) queues the ray and its contribution, to be added unless the
// predicates `owed` names (air_light_visibility_split) block it.
struct NoDefer {
    static constexpr bool on = false;
    SG_MEMBER void push(f3, f3, float, f3, uint32_t) {}
};
// The gradient structs have internal linkage, like the free functions their
// members call.
namespace {
// Adds weighted derivatives into a GRAD_COUNT accumulator (workgroup memory
// on the GPU). Every method takes d(loss) already multiplied out.
struct GradSink {
    static constexpr bool on = true;
    static constexpr bool capture = false;
    float *g;
    // Lane-private accumulators, slot k at dens[k * dstride]: the density
    // profile bins, then the five air coefficients (every contribution in
    // the air touches those, so they are summed per lane and flushed once).
    float *dens;
    uint32_t dstride;
    SG_MEMBER void add(uint32_t i, float v) const {
        if (v == 0.0f) return;
#if defined(__HIP_DEVICE_COMPILE__)
        atomicAdd(&g[i], v);
#else
        g[i] += v;
#endif
    }
    SG_MEMBER void albedo(uint32_t body, f3 d) const {
        add(GRAD_ALBEDO + 3u * body, d.x);
        add(GRAD_ALBEDO + 3u * body + 1u, d.y);
        add(GRAD_ALBEDO + 3u * body + 2u, d.z);
    }
    SG_MEMBER void ring(uint32_t bin, float d) const { add(GRAD_RING + bin, d); }
    SG_MEMBER void air(uint32_t k, float v) const { dens[(SG_DENSITY_BINS + k) * dstride] += v; }
    // d / d(beta_ext) per component; beta_ext = beta_rayleigh + beta_mie.
    SG_MEMBER void extinction(f3 d) const {
        air(0u, d.x);
        air(1u, d.y);
        air(2u, d.z);
        air(3u, d.x + d.y + d.z);
    }
    // A weighted contribution proportional to beta_phase = beta_r * rayleigh
    // + beta_m * mie(g): its derivatives through the phase coefficients.
    SG_MEMBER void phase(f3 wc, f3 beta_phase, float rayleigh, float mie, float dmie_dg, float beta_m) const {
        f3 q = f3{beta_phase.x > 0.0f ? wc.x / beta_phase.x : 0.0f, beta_phase.y > 0.0f ? wc.y / beta_phase.y : 0.0f,
                  beta_phase.z > 0.0f ? wc.z / beta_phase.z : 0.0f};
        air(0u, q.x * rayleigh);
        air(1u, q.y * rayleigh);
        air(2u, q.z * rayleigh);
        float sum = q.x + q.y + q.z;
        air(3u, sum * mie);
        air(4u, sum * beta_m * dmie_dg);
    }
    // d / d(profile bin) of a factor exp(-beta_ext * column) for weight
    // e = d(loss)/d log(factor) * (-beta_ext): each sample adds dt.
    SG_MEMBER void column(const Scene &sc, const ColumnBins &c, float e) const {
        if (!sc.density_profile || e == 0.0f) return;
        for (uint32_t k = 0; k < c.count; ++k) dens[(uint32_t)((c.bins >> (4u * k)) & 15u) * dstride] += e * c.dt;
    }
    SG_MEMBER void density(uint32_t bin, float v) const { dens[bin * dstride] += v; }
    // Move the lane's private accumulators into the gradient.
    SG_MEMBER void flush_private() const {
        for (uint32_t b = 0; b < SG_DENSITY_BINS; ++b) {
            add(GRAD_DENSITY + b, dens[b * dstride]);
            dens[b * dstride] = 0.0f;
        }
        for (uint32_t k = 0; k < 5u; ++k) {
            add(GRAD_AIR + k, dens[(SG_DENSITY_BINS + k) * dstride]);
            dens[(SG_DENSITY_BINS + k) * dstride] = 0.0f;
        }
    }
    // A weighted contribution behind a shadow ray that may cross the
    // tabulated ring.
    SG_MEMBER void shadow(const Scene &sc, const Ray &ray, uint32_t light, float weighted) const {
        uint32_t bin;
        float dlog;
        if (weighted != 0.0f && ring_profile_crossing(sc, ray, light, bin, dlog)) ring(bin, weighted * dlog);
    }
};
}  // namespace

// --- ground materials, themes 0-3 and 5 (src/shaders/ground.wgsl:256-666) ----
SG_FN float footprint_of(const Scene &sc, float distance) {
    return distance * 2.0f * sc.s.globals.cam_forward[3] / fmaxf(sc.s.globals.viewport[1], 1.0f);
}
SG_FN f3 forest_floor(const Scene &sc, f3 p, float distance, f3 normal) {
    float broad = vnoise(p * 0.045f + f3{7.0f, 0.0f, 13.0f});
    float middle = vnoise(p * 0.7f);
    float grain = vnoise(p * 18.0f);
    float detail = 1.0f - smoothstepf(0.006f, 0.055f, footprint_of(sc, distance));
    f3 moss = mix3(f3{0.028f, 0.054f, 0.014f}, f3{0.105f, 0.14f, 0.039f}, middle);
    f3 earth = mix3(f3{0.055f, 0.029f, 0.013f}, f3{0.17f, 0.115f, 0.057f}, middle);
    f3 color = mix3(earth, moss, smoothstepf(0.35f, 0.63f, broad));
    float needles = powf(fmaxf(0.0f, sinf(p.x * 92.0f + p.z * 137.0f + middle * 7.0f)), 18.0f) *
                    smoothstepf(0.50f, 0.68f, grain);
    color = mix3(color, f3{0.22f, 0.145f, 0.067f}, needles * detail * 0.5f);
    color = color * (1.0f + (grain - 0.5f) * 0.38f * detail);
    float slope = 1.0f - fmaxf(normal.y, 0.0f);
    f3 stone = mix3(f3{0.095f, 0.105f, 0.10f}, f3{0.24f, 0.23f, 0.20f}, middle);
    color = mix3(color, stone, smoothstepf(0.10f, 0.43f, slope + (broad - 0.5f) * 0.17f));
    float center = 2.0f + 0.30f * p.z + 2.0f * sinf(p.z * 0.055f);
    float edge = fabsf(p.x - center) + (middle - 0.5f) * 0.55f;
    float path = (1.0f - smoothstepf(0.8f, 2.1f, edge)) * (1.0f - smoothstepf(45.0f, 95.0f, p.z)) *
                 smoothstepf(-10.0f, 0.0f, p.z);
    f3 gravel = mix3(f3{0.12f, 0.095f, 0.062f}, f3{0.26f, 0.23f, 0.17f}, grain * detail + 0.5f * (1.0f - detail));
    return mix3(color, gravel, path * 0.85f);
}
// Ground theme, the low byte of ground_counts.z (src/shaders/terrain.wgsl:22).
SG_FN uint32_t ground_theme(const Scene &sc) { return sc.s.globals.ground_counts[2] & 255u; }
// Theme 1, desert: sand and rock by slope, gravel grain, a meandering
// two-rut trail (src/shaders/ground.wgsl:281-304). Out of line, like the
// solid-body styles, to keep it out of the forest path's registers.
SG_OUTLINE f3 desert_floor(float footprint, f3 p, f3 normal) {
    const float broad = vnoise(p * 0.007f + f3{19.0f, 2.0f, 4.0f});
    const float gravel = vnoise(p * 2.8f);
    const float detail = 1.0f - smoothstepf(0.025f, 0.2f, footprint);
    const f3 sand = mix3(f3{0.31f, 0.21f, 0.125f}, f3{0.53f, 0.40f, 0.26f}, broad);
    const f3 rock = mix3(f3{0.18f, 0.10f, 0.055f}, f3{0.38f, 0.245f, 0.145f}, broad);
    f3 color = mix3(sand, rock, smoothstepf(0.08f, 0.4f, 1.0f - fmaxf(normal.y, 0.0f)));
    color = color * (1.0f + (gravel - 0.5f) * 0.32f * detail);
    const float center = 2.8f + 0.55f * sinf(p.z / 38.0f) + 1.4f * sinf(p.z / 210.0f);
    const float edge = fabsf(p.x - center);
    const float aa = fmaxf(0.015f, footprint * 0.5f);
    const float fade = (1.0f - smoothstepf(0.25f, 0.9f, footprint)) * (1.0f - smoothstepf(1000.0f, 2200.0f, fabsf(p.z))) *
                       (1.0f - smoothstepf(0.12f, 0.35f, 1.0f - fmaxf(normal.y, 0.0f)));
    const float worn = 1.0f - smoothstepf(0.8f, 1.2f, edge);
    const float rut_edge = fabsf(edge - 0.68f) + (gravel - 0.5f) * 0.1f * detail;
    const float tracks = 1.0f - smoothstepf(0.12f - aa, 0.23f + aa, rut_edge);
    const f3 packed = sand * (0.72f + 0.10f * gravel * detail);
    color = mix3(color, sand * 0.93f, worn * fade * 0.35f);
    return mix3(color, packed, tracks * fade * 0.8f);
}
// Theme 2, grey lunar regolith: broad mottling, brighter steep walls, grain
// (src/shaders/ground.wgsl:306-317).
SG_OUTLINE f3 crater_floor(float footprint, f3 p, f3 normal) {
    const float broad = vnoise(p * 0.004f + f3{3.0f, 0.0f, 8.0f});
    const float middle = vnoise(p * 0.09f);
    const float grain = vnoise(p * 3.1f);
    const float detail = 1.0f - smoothstepf(0.025f, 0.2f, footprint);
    float tone = mixf(0.11f, 0.21f, broad) * (0.86f + 0.28f * middle);
    tone *= 1.0f + 0.35f * smoothstepf(0.15f, 0.5f, 1.0f - fmaxf(normal.y, 0.0f));
    tone *= 1.0f + (grain - 0.5f) * 0.3f * detail;
    return f3{tone, tone * 0.97f, tone * 0.93f};
}
// Ground curvature below (px, pz) for a host of radius r (terrain.wgsl:32-36).
SG_FN float ground_curvature(float r, float px, float pz) {
    const float d2 = px * px + pz * pz;
    return -d2 / (r + sqrtf(fmaxf(r * r - d2, 0.0f)));
}
// The ice flows west-south-west; the lookout stands on a nunatak of height
// 180 m and base radius 700 m (ground.wgsl:318-331, terrain.rs).
constexpr float ICE_FLOW = 4.4f;
constexpr float DUNE_WIND = 0.61f;
constexpr float DUNE_PAN_LEVEL = 2.0f;
SG_FN float lookout_hill(float r) {
    const float u = fminf(sqrtf(r * r + 30.0f * 30.0f) / 700.0f, 1.0f);
    return 180.0f * (1.0f - u) * (1.0f - u);
}
// Gap between the nearest two jittered cells: polygonal cracks
// (ground.wgsl:490-507).
SG_FN float dune_cracks(float px, float py) {
    const float gx = floorf(px), gy = floorf(py);
    const float fx = px - gx, fy = py - gy;
    float d1 = 8.0f, d2 = 8.0f;
    for (int j = -1; j <= 1; ++j) {
        for (int i = -1; i <= 1; ++i) {
            const float ox = (float)i, oy = (float)j;
            const f3 h = hash33(f3{gx + ox, gy + oy, 3.0f});
            const float dx = ox + 0.15f + 0.7f * h.x - fx, dy = oy + 0.15f + 0.7f * h.y - fy;
            const float d = sqrtf(dx * dx + dy * dy);
            if (d < d1) {
                d2 = d1;
                d1 = d;
            } else if (d < d2) {
                d2 = d;
            }
        }
    }
    return d2 - d1;
}
// Crevasses across the ice flow: (coverage of the pixel, 0 at a crack's lip
// to 1 at its centre) (ground.wgsl:333-362).
SG_FN float2 ice_crevasse(float qx, float qz, float footprint) {
    const float r = sqrtf(qx * qx + qz * qz);
    const float tension = smoothstepf(650.0f, 900.0f, r) * (1.0f - smoothstepf(1800.0f, 2600.0f, r));
    const float strain = fmaxf(smoothstepf(0.47f, 0.58f, fbm(f3{qx / 2600.0f + 4.0f, 2.0f, qz / 2600.0f}, 2u)), tension) *
                         smoothstepf(600.0f, 750.0f, r);
    if (strain <= 0.0f) return make_float2(0.0f, 0.0f);
    const float flx = sinf(ICE_FLOW), flz = cosf(ICE_FLOW);
    const float along = qx * flx + qz * flz + 280.0f * (vnoise(f3{qx / 1500.0f, qz / 1500.0f, 1.0f}) - 0.5f);
    const float across = qx * flz - qz * flx + 120.0f * (vnoise(f3{qx / 900.0f, qz / 900.0f, 5.0f}) - 0.5f);
    const float spacing = 60.0f;
    const float row = floorf(along / spacing);
    const uint32_t key = hash_u32(((uint32_t)(int32_t)row * 0x8da6b343u) ^ 0xc3e5u);
    const float length_m = 120.0f + 300.0f * (float)(key & 255u) / 255.0f;
    const float s = across / length_m + (float)((key >> 8) & 255u) / 255.0f;
    const float segment = floorf(s);
    const float lens = powf(fmaxf(sinf(PI * (s - segment)), 0.0f), 0.7f);
    const uint32_t open = hash_u32(key ^ (uint32_t)(int32_t)segment);
    if ((open & 255u) < 70u) return make_float2(0.0f, 0.0f);
    const float centre = 0.5f + 0.3f * ((float)((open >> 8) & 255u) / 255.0f - 0.5f);
    const float half_width = 0.5f * (2.0f + 14.0f * (float)((open >> 16) & 255u) / 255.0f) * lens * strain;
    const float d = fabsf(along / spacing - row - centre) * spacing + 0.8f * (vnoise(f3{qx * 0.15f, qz * 0.15f, 9.0f}) - 0.5f);
    const float aa = fmaxf(footprint, 1.0e-3f);
    const float coverage = clampf((half_width - d) / aa + 0.5f, 0.0f, 1.0f) * fminf(1.0f, 2.0f * half_width / aa);
    return make_float2(coverage, 1.0f - clampf(d / fmaxf(half_width, 1.0e-3f), 0.0f, 1.0f));
}
// Theme 3, Patagonian ice field: snow, bare blue ice with fractures,
// crevasses, debris, granite spires with lichen and snow ledges
// (ground.wgsl:365-428). `footprint` is the raw pixel footprint in metres.
SG_OUTLINE f3 alpine_floor(float footprint_m, float scale, float radius, f3 p, f3 normal) {
    f3 q = f3{p.x, p.y - ground_curvature(radius, p.x, p.z), p.z} / scale;
    const float hill = lookout_hill(sqrtf(q.x * q.x + q.z * q.z));
    q.y -= hill;
    const float footprint = footprint_m / scale;
    const float detail = 1.0f - smoothstepf(0.025f, 0.2f, footprint);
    const float slope = 1.0f - fmaxf(normal.y, 0.0f);
    const float broad = vnoise(q * 0.004f + f3{11.0f, 4.0f, 1.0f});
    const float middle = vnoise(q * 0.2f);
    const float grain = vnoise(q * 4.0f);
    const float flx = sinf(ICE_FLOW), flz = cosf(ICE_FLOW);
    const float along = q.x * flx + q.z * flz;
    const float across = q.x * flz - q.z * flx;
    const float crust = smoothstepf(0.45f, 0.6f, fbm(f3{along * 0.004f, 2.0f, across * 0.012f}, 3u));
    f3 snow = mix3(f3{0.86f, 0.88f, 0.91f}, f3{0.76f, 0.8f, 0.86f}, crust);
    snow = snow * (0.97f + 0.05f * middle);
    const float bare = smoothstepf(0.47f, 0.53f, fbm(q * 0.0016f + f3{3.0f, 0.0f, 8.0f}, 4u) + 0.1f * (broad - 0.5f));
    f3 ice = mix3(f3{0.36f, 0.55f, 0.7f}, f3{0.52f, 0.67f, 0.78f}, middle);
    const float fracture_width = 0.03f + footprint / 4.0f;
    const float fractures =
        (1.0f - smoothstepf(0.0f, fracture_width, dune_cracks(q.x / 4.0f + 9.0f, q.z / 4.0f + 9.0f))) * 0.03f / fracture_width;
    ice = mix3(ice, f3{0.82f, 0.86f, 0.9f}, fminf(fractures, 1.0f) * 0.7f);
    ice = ice * (1.0f + (grain - 0.5f) * 0.12f * detail);
    f3 glacier = mix3(snow, ice, bare);
    const f3 debris = mix3(f3{0.12f, 0.115f, 0.11f}, f3{0.26f, 0.24f, 0.22f}, grain);
    const float dirty = smoothstepf(6.0f, 30.0f, q.y) * (1.0f - smoothstepf(30.0f, 70.0f, q.y));
    glacier = mix3(glacier, mix3(debris, glacier, 0.45f), dirty * smoothstepf(0.4f, 0.7f, broad) * 0.8f);
    const float wall = smoothstepf(0.35f, 0.6f, slope) * (1.0f - smoothstepf(20.0f, 45.0f, q.y));
    const float depth = smoothstepf(4.0f, -4.0f, q.y);
    const f3 blue = mix3(f3{0.32f, 0.55f, 0.72f}, f3{0.1f, 0.25f, 0.42f}, depth);
    glacier = mix3(glacier, blue * (0.9f + 0.2f * middle), wall);
    const float2 cr = ice_crevasse(q.x, q.z, footprint);
    const float crack_x = cr.x * (1.0f - smoothstepf(10.0f, 25.0f, q.y));
    const float crack_y = cr.y * (1.0f - smoothstepf(10.0f, 25.0f, q.y));
    const float lip = (1.0f - smoothstepf(0.0f, 0.12f, crack_y)) * crack_x;
    const f3 inside = mix3(f3{0.3f, 0.55f, 0.74f}, f3{0.02f, 0.06f, 0.12f}, smoothstepf(0.1f, 0.8f, crack_y));
    glacier = mix3(glacier, inside, crack_x);
    glacier = mix3(glacier, f3{0.9f, 0.92f, 0.95f}, lip * 0.5f);
    const float mountain = fmaxf(smoothstepf(25.0f, 70.0f, q.y + 30.0f * (broad - 0.5f)), smoothstepf(4.0f, 15.0f, hill));
    const float joints =
        1.0f - smoothstepf(0.0f, 0.06f + footprint / 3.0f, dune_cracks((q.x + q.y * 0.5f) / 3.0f, q.z / 3.0f));
    f3 granite = mix3(f3{0.09f, 0.085f, 0.08f}, f3{0.24f, 0.215f, 0.19f}, broad);
    granite = granite * (0.85f + 0.3f * middle - 0.35f * joints * (1.0f - smoothstepf(0.1f, 1.0f, footprint)));
    const float lichen = smoothstepf(0.68f, 0.78f, vnoise(q * 0.9f)) * (1.0f - smoothstepf(0.6f, 0.9f, slope));
    granite = mix3(granite, f3{0.3f, 0.17f, 0.06f}, lichen * 0.5f * detail);
    const float cover = 1.0f - smoothstepf(0.17f, 0.33f, slope + (broad - 0.5f) * 0.18f + (middle - 0.5f) * 0.08f);
    const float scoured = smoothstepf(0.52f, 0.62f, fbm(q * 0.012f + f3{2.0f, 0.0f, 6.0f}, 4u)) *
                          smoothstepf(25.0f, 90.0f, hill) * (1.0f - smoothstepf(150.0f, 172.0f, hill));
    const float cover_hill = cover * (1.0f - scoured);
    const f3 peak = mix3(granite, snow, cover_hill);
    f3 color = mix3(glacier, peak, mountain);
    return color * (1.0f + (grain - 0.5f) * 0.08f * detail);
}
// Sastrugi: the ground-frame (x, z) gradient of wind-carved snow ridges
// (ground.wgsl:430-461); the caller adds it to the normal.
SG_OUTLINE float2 alpine_gradient(float footprint, f3 p, f3 n) {
    const float snowy = 1.0f - smoothstepf(0.3f, 0.5f, 1.0f - n.y);
    const float flx = sinf(ICE_FLOW), flz = cosf(ICE_FLOW);
    const float sx = flz, sz = -flx;
    float gx = 0.0f, gz = 0.0f;
    float scale = 1.2f, strength = 0.2f;
    for (uint32_t octave = 0; octave < 3u; ++octave) {
        const float fade = 1.0f - smoothstepf(0.08f * scale, 0.25f * scale, footprint);
        if (fade > 0.0f) {
            const float a = (p.x * flx + p.z * flz) / (6.0f * scale);
            const float c = (p.x * sx + p.z * sz) / scale;
            const float e = 0.08f;
            const float o = (float)octave * 13.0f;
            const float da = vnoise(f3{a + e, o, c}) - vnoise(f3{a - e, o, c});
            const float dc = vnoise(f3{a, o, c + e}) - vnoise(f3{a, o, c - e});
            const float crest = 1.0f + 1.5f * (1.0f - fabsf(vnoise(f3{a, o, c}) * 2.0f - 1.0f));
            const float w = strength * fade * crest * scale;
            gx += (flx * da / (6.0f * scale) + sx * dc / scale) / (2.0f * e) * w;
            gz += (flz * da / (6.0f * scale) + sz * dc / scale) / (2.0f * e) * w;
        }
        scale *= 6.0f;
        strength *= 0.6f;
    }
    return make_float2(gx * snowy, gz * snowy);
}
// Theme 5, Namib sand: dune plinths and crests, slip faces, mineral lag,
// gravel corridors and cracked clay pans (ground.wgsl:483-569).
SG_OUTLINE f3 dune_floor(float footprint_m, float scale, float radius, f3 p, f3 normal) {
    const f3 q = f3{p.x, p.y - ground_curvature(radius, p.x, p.z), p.z} / scale;
    const float footprint = footprint_m / scale;
    const float detail = 1.0f - smoothstepf(0.02f, 0.15f, footprint);
    const float slope = 1.0f - fmaxf(normal.y, 0.0f);
    const float broad = vnoise(q * 0.0015f + f3{1.0f, 7.0f, 3.0f});
    const float middle = vnoise(q * 0.05f);
    const float rise = smoothstepf(6.0f, 160.0f, q.y + 40.0f * (broad - 0.5f));
    const f3 plinth = mix3(f3{0.3f, 0.095f, 0.038f}, f3{0.36f, 0.12f, 0.045f}, broad);
    const f3 upper = mix3(f3{0.42f, 0.155f, 0.06f}, f3{0.47f, 0.19f, 0.072f}, broad);
    f3 sand = mix3(plinth, upper, rise) * (0.94f + 0.08f * middle);
    const float flat = sqrtf(normal.x * normal.x + normal.z * normal.z);
    if (flat > 1.0e-3f) {
        const float dx = -normal.x / flat, dz = -normal.z / flat;
        const float across = q.x * -dz + q.z * dx;
        const float along = q.x * dx + q.z * dz;
        const float tongues = fbm(f3{across * 0.18f, along * 0.012f, 5.0f}, 3u);
        const float slip = smoothstepf(0.32f, 0.45f, slope);
        sand = mix3(sand, sand * f3{1.06f, 1.12f, 1.16f} * (0.86f + 0.28f * tongues), slip);
    }
    const float lag = smoothstepf(0.6f, 0.76f, fbm(q * f3{0.013f, 0.0f, 0.004f} + f3{9.0f, 0.0f, 2.0f}, 4u)) *
                      (1.0f - smoothstepf(0.04f, 0.16f, slope)) * (1.0f - smoothstepf(10.0f, 45.0f, q.y));
    sand = mix3(sand, f3{0.13f, 0.065f, 0.05f}, lag * 0.55f);
    const float corridor =
        (1.0f - smoothstepf(5.0f, 16.0f, q.y + 6.0f * (broad - 0.5f))) * (1.0f - smoothstepf(0.03f, 0.1f, slope));
    const f3 gravel = mix3(f3{0.39f, 0.21f, 0.115f}, f3{0.45f, 0.28f, 0.16f}, vnoise(q * 0.3f));
    const float pebbles = smoothstepf(0.72f, 0.8f, vnoise(q * 7.0f)) * detail;
    sand = mix3(sand, gravel * (1.0f - 0.45f * pebbles), corridor * 0.6f);
    const float pans[2][3] = {{842.6f, 707.1f, 560.0f}, {4346.7f, 1164.7f, 900.0f}};
    float pan = 0.0f;
    for (uint32_t i = 0; i < 2u; ++i) {
        const float ex = q.x - pans[i][0], ez = q.z - pans[i][1];
        pan = fmaxf(pan, 1.0f - smoothstepf(pans[i][2] * 1.2f, pans[i][2] * 1.35f + 140.0f, sqrtf(ex * ex + ez * ez)));
    }
    pan *= 1.0f - smoothstepf(DUNE_PAN_LEVEL + 0.25f, DUNE_PAN_LEVEL + 1.5f, q.y);
    if (pan > 0.0f) {
        const float tone = fbm(q * 0.012f + f3{4.0f, 0.0f, 1.0f}, 4u);
        f3 clay = mix3(f3{0.4f, 0.35f, 0.29f}, f3{0.6f, 0.56f, 0.49f}, smoothstepf(0.3f, 0.7f, tone));
        clay = clay * (0.82f + 0.18f * smoothstepf(0.35f, 1.0f, pan));
        const float fine_width = 0.04f + footprint / 0.9f;
        const float cracks = (1.0f - smoothstepf(0.0f, fine_width, dune_cracks(q.x / 0.9f, q.z / 0.9f))) * 0.04f / fine_width;
        const float wide_width = 0.03f + footprint / 6.0f;
        const float large =
            (1.0f - smoothstepf(0.0f, wide_width, dune_cracks(q.x / 6.0f + 17.0f, q.z / 6.0f + 17.0f))) * 0.03f / wide_width;
        clay = clay * (1.0f - 0.6f * fminf(cracks, 1.0f) - 0.45f * fminf(large, 1.0f));
        clay = clay * (0.92f + 0.16f * vnoise(q * 2.3f) * (1.0f - smoothstepf(0.04f, 0.25f, footprint)));
        sand = mix3(sand, clay, pan);
    }
    return sand * (1.0f + (vnoise(q * 9.0f) - 0.5f) * 0.12f * detail);
}
// Wind ripples across the wind (ground.wgsl:571-581): the (x, z) gradient.
SG_FN float2 dune_ripples(f3 p, float wavelength, float warp, float amplitude) {
    const float wx = sinf(DUNE_WIND), wz = cosf(DUNE_WIND);
    const float sx = wz, sz = -wx;
    const float bend = warp * (vnoise(f3{p.x / (9.0f * wavelength), p.z / (9.0f * wavelength), 1.0f}) +
                               0.35f * vnoise(f3{p.x / (2.5f * wavelength), p.z / (2.5f * wavelength), 7.0f}));
    const float phase = (p.x * wx + p.z * wz) / wavelength + bend + 0.4f * sinf((p.x * sx + p.z * sz) / (3.0f * wavelength));
    const float theta = TAU * phase;
    const float slope = cosf(theta + 0.65f * sinf(theta)) * (1.0f + 0.65f * cosf(theta));
    const float patches = smoothstepf(0.25f, 0.6f, vnoise(f3{p.x / (14.0f * wavelength), p.z / (14.0f * wavelength), 3.0f}));
    const float k = slope * amplitude * TAU / wavelength * patches;
    return make_float2(wx * k, wz * k);
}
// Ripple gradient of the dunes (ground.wgsl:583-595).
SG_OUTLINE float2 dune_gradient(float footprint, float scale, float radius, f3 p, f3 n) {
    const float gentle = 1.0f - smoothstepf(0.22f, 0.38f, 1.0f - n.y);
    const float qy = (p.y - ground_curvature(radius, p.x, p.z)) / scale;
    const float dry = smoothstepf(DUNE_PAN_LEVEL + 0.4f, DUNE_PAN_LEVEL + 1.5f, qy);
    float gx = 0.0f, gz = 0.0f;
    const float fine = 1.0f - smoothstepf(0.03f, 0.07f, footprint);
    if (fine > 0.0f) {
        const float2 g = dune_ripples(p, 0.12f, 2.5f, 0.006f);
        gx += g.x * fine;
        gz += g.y * fine;
    }
    const float coarse = (1.0f - smoothstepf(0.15f, 0.4f, footprint)) * (1.0f - smoothstepf(9.0f, 16.0f, qy));
    if (coarse > 0.0f) {
        const float2 g = dune_ripples(p + f3s(31.0f), 0.8f, 2.0f, 0.03f);
        gx += g.x * coarse;
        gz += g.y * coarse;
    }
    return make_float2(gx * gentle * dry, gz * gentle * dry);
}
SG_FN d3 landscape_point(const Scene &sc, const Ray &ray, const Hit &hit) {
    return ground_origin(sc, ray) + tod(to_ground(sc, ray.direction)) * ((double)hit.distance * METRES_PER_AU_D);
}
// Ground material at surface-frame point p (metres) of landscape element
// `index`, seen from `distance` metres, with surface-frame normal n.
SG_FN f3 ground_albedo(const Scene &sc, f3 p, float distance, f3 n, uint32_t index) {
    if (index == SG_LOCAL_BASE || index == SG_TERRAIN_ID) {
        switch (ground_theme(sc)) {
        case 1u: return desert_floor(footprint_of(sc, distance), p, n);
        case 2u: return crater_floor(footprint_of(sc, distance), p, n);
        case 3u: return alpine_floor(footprint_of(sc, distance), terrain_scale(sc), sc.s.globals.ground_up[3], p, n);
        case 5u: return dune_floor(footprint_of(sc, distance), terrain_scale(sc), sc.s.globals.ground_up[3], p, n);
        default: return forest_floor(sc, p, distance, n);
        }
    }
    const sg_primitive &object = sc.objects[index - SG_LOCAL_BASE - 1u];
    f3 tint = {object.albedo[0], object.albedo[1], object.albedo[2]};
    float kind = object.albedo[3];
    if (kind == 0.0f) return tint;
    float middle = vnoise(p * 1.7f);
    float fine = mixf(vnoise(p * 12.0f), 0.5f, smoothstepf(18.0f, 110.0f, distance));
    if (kind > 4.5f) return tint * (0.65f + 0.6f * middle);
    if (kind > 2.5f) {
        f3 q = rotate_north(p - f3{object.center[0], object.center[1], object.center[2]}, -object.extent[3]) /
               f3{object.extent[0], object.extent[1], object.extent[2]};
        if (kind > 3.5f || (q.y < -0.35f && sqrtf(q.x * q.x + q.z * q.z) < 0.09f)) {
            float bark = vnoise(p * f3{25.0f, 1.5f, 25.0f});
            return f3{0.13f, 0.073f, 0.032f} * (0.5f + 0.85f * bark);
        }
        return tint * (0.7f + 0.55f * middle + 0.2f * fine);
    }
    float strata = sinf(p.y * 8.0f + 2.5f * middle);
    float lichen = smoothstepf(0.52f, 0.76f, middle) * fmaxf(n.y, 0.0f);
    f3 stone = tint * (0.7f + 0.45f * fine + 0.07f * strata);
    if (ground_theme(sc) != 0u) return stone;  // only the forest's rocks carry moss and lichen
    return mix3(stone, f3{0.11f, 0.145f, 0.037f}, lichen * 0.65f);
}
SG_FN f3 landscape_albedo(const Scene &sc, const Ray &ray, const Hit &hit) {
    return ground_albedo(sc, tof(landscape_point(sc, ray, hit)), hit.distance * METRES_PER_AU,
                         to_ground(sc, hit.normal), hit.index);
}
// Shading normal (world frame) at surface-frame point p of landscape element
// `index`, seen from `distance` metres, with world-frame geometric normal.
SG_FN f3 ground_shading_normal(const Scene &sc, f3 p, float distance, f3 normal, uint32_t index) {
    if (index != SG_LOCAL_BASE && index != SG_TERRAIN_ID) return normal;
    const uint32_t theme = ground_theme(sc);
    if (theme == 3u || theme == 5u) {
        const f3 n = to_ground(sc, normal);
        const float2 g = theme == 3u ? alpine_gradient(footprint_of(sc, distance), p, n)
                                     : dune_gradient(footprint_of(sc, distance), terrain_scale(sc),
                                                     sc.s.globals.ground_up[3], p, n);
        return normalize(normal + from_ground(sc, f3{-g.x, 0.0f, -g.y}));
    }
    float strength = 0.11f * (1.0f - smoothstepf(8.0f, 65.0f, distance));
    float dx = vnoise((p + f3{0.025f, 0.f, 0.f}) * 6.0f) - vnoise((p - f3{0.025f, 0.f, 0.f}) * 6.0f);
    float dz = vnoise((p + f3{0.f, 0.f, 0.025f}) * 6.0f) - vnoise((p - f3{0.f, 0.f, 0.025f}) * 6.0f);
    return normalize(normal + from_ground(sc, f3{-dx, 0.0f, -dz}) * strength);
}
SG_FN f3 landscape_shading_normal(const Scene &sc, const Ray &ray, const Hit &hit) {
    if (hit.index != SG_LOCAL_BASE && hit.index != SG_TERRAIN_ID) return hit.normal;
    return ground_shading_normal(sc, tof(landscape_point(sc, ray, hit)), hit.distance * METRES_PER_AU, hit.normal,
                                 hit.index);
}
// 5 mm outgoing offset with reprojection onto the host sphere, FP64.
SG_FN Ray landscape_outgoing(const Scene &sc, const Ray &ray, const Hit &hit) {
    d3 p = landscape_point(sc, ray, hit);
    if (hit.index == SG_LOCAL_BASE) {
        double radius = sc.s.globals.ground_up[3];
        double d2 = p.x * p.x + p.z * p.z;
        if (p.y > -radius && d2 < radius * radius) p.y = -d2 / (radius + sqrt(radius * radius - d2));
    }
    p = p + tod(to_ground(sc, hit.normal)) * 0.005;
    p.y -= (double)sc.s.globals.ground_east[3];
    return Ray{from_ground_d(sc, p) / METRES_PER_AU_D, hit.normal, SG_MISS};
}

// --- host atmosphere (src/shaders/pathtrace.wgsl:927-1234) ------------------
struct Scatter {
    f3 transmittance, inscatter;
};
SG_FN f3 atmo_center(const Scene &sc) { const float *v = sc.s.globals.atmo_center; return f3{v[0], v[1], v[2]}; }
SG_FN float host_radius(const Scene &sc) { return sc.s.globals.atmo_center[3]; }
SG_FN float atmosphere_top(const Scene &sc) { return sc.s.globals.atmo_center[3] + sc.s.globals.atmo_params[2]; }
SG_FN float atmo_scale_height(const Scene &sc) { return sc.s.globals.atmo_params[1]; }
SG_FN f3 beta_rayleigh(const Scene &sc) { const float *v = sc.s.globals.atmo_rayleigh; return f3{v[0], v[1], v[2]}; }
SG_FN float beta_mie(const Scene &sc) { return sc.s.globals.atmo_rayleigh[3]; }

SG_FN float shell_exit(const Scene &sc, f3 origin, f3 dir) {
    f3 oc = origin - atmo_center(sc);
    float b = dot(oc, dir);
    float top = atmosphere_top(sc);
    float c = dot(oc, oc) - top * top;
    float disc = b * b - c;
    if (disc < 0.0f) return 0.0f;
    return -b + sqrtf(disc);
}
// Whether the Mie density has its own scale height (an extension: the
// reference shader uses one exp(-h / H) for both species).
SG_FN bool two_heights(const Scene &sc) { return sc.mie_height > 0.0f && !sc.density_profile; }
// Midpoint column of the density along `dir` to the top of the air
// (src/shaders/pathtrace.wgsl:958-971). With two scale heights `mie`
// receives the Mie column; with one it is left alone (the columns are equal).
SG_FN float density_integral(const Scene &sc, f3 origin, f3 dir, uint32_t steps, ColumnBins *trace = nullptr,
                             float *mie = nullptr) {
    float distance = shell_exit(sc, origin, dir);
    if (distance <= 0.0f) {
        if (mie) *mie = 0.0f;
        return 0.0f;
    }
    float dt = distance / (float)steps;
    float inv_h = 1.0f / atmo_scale_height(sc);
    const bool split = mie && two_heights(sc);
    float inv_hm = split ? 1.0f / sc.mie_height : 0.0f;
    float r = host_radius(sc);
    f3 c = atmo_center(sc);
    float sum = 0.0f, sum_m = 0.0f;
    for (uint32_t i = 0; i < steps; ++i) {
        SG_COUNT(SG_EV_DENSITY_STEP);
        f3 p = origin + dir * (((float)i + 0.5f) * dt);
        float h = fmaxf(length(p - c) - r, 0.0f);
        if (sc.density_profile) {
            const uint32_t bin = density_bin(sc, h);
            sum += sc.density_profile[bin];
            if (trace && i < COLUMN_BIN_SLOTS) trace->bins |= (uint64_t)bin << (4u * i);
        } else {
            sum += sg_exp(-h * inv_h);
            if (split) sum_m += sg_exp(-h * inv_hm);
        }
    }
    if (trace) {
        trace->dt = dt;
        trace->count = steps;
        trace->total = sum * dt;
    }
    if (split) *mie = sum_m * dt;
    return sum * dt;
}
// FP64 rebasing of an anchored origin to the host centre.
SG_FN d3 host_offset(const Scene &sc, const Ray &origin) {
    d3 c = tod(atmo_center(sc));
    if (origin.anchor == SG_MISS) return origin.offset - c;
    return (body_centre(sc, origin.anchor) - c) + origin.offset;
}
SG_FN bool inside_air(const Scene &sc, const Ray &origin) {
    if (sc.s.globals.atmo_params[3] < 0.5f) return false;
    f3 oc = tof(host_offset(sc, origin));
    float top = atmosphere_top(sc);
    return dot(oc, oc) < top * top;
}
// Optical depth of a Rayleigh column and a Mie column. With one scale
// height both are the same column and the depth is (beta_r + beta_m) *
// column, as the reference writes it.
SG_FN f3 air_depth(const Scene &sc, float column, float mie_column) {
    if (two_heights(sc)) return beta_rayleigh(sc) * column + f3s(beta_mie(sc) * mie_column);
    return (beta_rayleigh(sc) + f3s(beta_mie(sc))) * column;
}
// Density column from an anchored origin to the top of the air; negative
// when the origin is outside the air (no extinction). `mie` receives the Mie
// column under two scale heights.
SG_FN float air_column(const Scene &sc, const Ray &origin, f3 dir, ColumnBins *trace = nullptr,
                       float *mie = nullptr) {
    if (!inside_air(sc, origin)) return -1.0f;
    f3 p = atmo_center(sc) + tof(host_offset(sc, origin));
    return density_integral(sc, p, dir, sc.air_sun_steps, trace, mie);
}
SG_FN f3 air_extinction(const Scene &sc, float column, float mie_column) {
    if (column < 0.0f) return f3s(1.0f);
    return exp3(-air_depth(sc, column, mie_column));
}
SG_FN f3 air_transmittance(const Scene &sc, const Ray &origin, f3 dir) {
    float mie = 0.0f;
    float column = air_column(sc, origin, dir, nullptr, &mie);
    return air_extinction(sc, column, mie);
}
SG_FN float rayleigh_phase(float c) { return 3.0f / (16.0f * PI) * (1.0f + c * c); }
SG_FN float mie_phase(float c, float g) {
    float g2 = g * g;
    // u^1.5 as u * sqrt(u): the WGSL's pow(u, 1.5) without the general
    // power's log and exp (4% of the camera-air step).
    const float u = 1.0f + g2 - 2.0f * g * c;
    return (1.0f - g2) / (4.0f * PI * (u * sqrtf(u)));
}
SG_FN float mie_phase_dg(float c, float g) {
    float u = 1.0f + g * g - 2.0f * g * c;
    float s = sqrtf(u);
    return (-2.0f * g / (u * s) - 3.0f * (1.0f - g * g) * (g - c) / (u * u * s)) / (4.0f * PI);
}
// d log(segment) / d(beta_ext) for the segment integral of one view step.
SG_FN float segment_dlog(float tau, float column, float segment, float beta_ext) {
    if (segment <= 0.0f) return 0.0f;
    if (tau < 1.0e-3f) return -0.5f * column * column / segment;
    return (column * expf(-tau) / segment - 1.0f) / fmaxf(beta_ext, 1.0e-20f);
}

struct ReflectorSample {
    uint32_t index;
    float inverse_probability;
};
SG_FN Ray camera_ray(f3 origin, f3 dir) { return Ray{tod(origin), dir, SG_MISS}; }

// src/shaders/pathtrace.wgsl:1021-1043
SG_FN ReflectorSample sample_reflector(const Scene &sc, f3 origin, uint32_t &rng) {
    Ray ray = camera_ray(origin, f3{0.f, 0.f, -1.f});
    uint32_t host = (uint32_t)sc.s.globals.atmo_params[3] - 1u;
    uint32_t selected = SG_MISS;
    float selected_weight = 0.0f, total = 0.0f;
    for (uint32_t b = 0; b < body_count(sc); ++b) {
        const Body &bd = body(sc, b);
        if (b == host || bd.material.w > 0.5f) continue;
        float weight = cone_width(sc, ray, b) * maxc(xyz(bd.material));
        if (weight <= 0.0f) continue;
        total += weight;
        if (rnd(rng) * total < weight) {
            selected = b;
            selected_weight = weight;
        }
    }
    return ReflectorSample{selected, total / fmaxf(selected_weight, 1.0e-30f)};
}
// src/shaders/pathtrace.wgsl:1048-1071
SG_FN f3 reflected_radiance(const Scene &sc, const Ray &ray, uint32_t index, uint32_t &rng) {
    Hit hit = sphere_hit(sc, ray, index);
    if (hit.index == SG_MISS) return f3s(0.0f);
    const Body &b = body(sc, index);
    f3 albedo = surface_albedo(b, hit.normal);
    Ray outgoing{tod(hit.normal * (b.center.w * (1.0f + 2.0e-6f))), hit.normal, index};
    f3 radiance = f3s(0.0f);
    for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
        uint32_t s = light_index(sc, slot);
        if (body(sc, s).material.w < 0.5f) continue;
        LightSample light = sample_light(sc, outgoing, s, rng);
        float cosine = fmaxf(dot(hit.normal, light.direction), 0.0f);
        if (cosine > 0.0f) {
            outgoing.direction = light.direction;
            radiance += albedo * xyz(body(sc, s).material) *
                        (cosine * light_visibility(sc, outgoing, s) / (PI * light.pdf));
        }
    }
    return radiance;
}
// src/shaders/pathtrace.wgsl:1076-1087
SG_FN float brightest_risen_light(const Scene &sc, f3 origin) {
    Ray ray = camera_ray(origin, f3{0.f, 0.f, -1.f});
    float brightest = 0.0f;
    for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
        uint32_t b = light_index(sc, slot);
        if (body(sc, b).material.w < 0.5f) continue;
        if (dot(normalize(tof(relative_center(sc, ray, b))), g_up(sc)) <= 0.0f) continue;
        brightest = fmaxf(brightest, maxc(xyz(body(sc, b).material)) * TAU * cone_width(sc, ray, b));
    }
    return brightest;
}

// Single scattering along a camera ray (src/shaders/pathtrace.wgsl:1092-1213).
// With a gradient sink, `wt` is d(loss)/d(inscatter) and the derivatives of
// the in-scatter with respect to the air coefficients, the reflector's
// albedo and the tabulated ring profile (shadow rays) are added to `grad`.
template <class G = NoGrad, class D = NoDefer>
SG_FN Scatter atmosphere(const Scene &sc, f3 origin, f3 dir, float limit, uint32_t &rng, uint32_t accumulated,
                         const G &grad = G(), f3 wt = f3s(0.0f), D *defer = nullptr) {
    Scatter result{f3s(1.0f), f3s(0.0f)};
    if (sc.s.globals.atmo_params[3] < 0.5f) return result;
    float exit = shell_exit(sc, origin, dir);
    float t_max = fminf(limit, exit);
    if (t_max <= 0.0f) return result;
    f3 beta_r = beta_rayleigh(sc);
    float beta_m = beta_mie(sc);
    f3 beta_ext = beta_r + f3s(beta_m);
    float g = sc.s.globals.atmo_params[0];
    float r = host_radius(sc);
    f3 center = atmo_center(sc);
    float H = atmo_scale_height(sc);
    const uint32_t view_steps = sc.air_view_steps;
    uint32_t steps = umin(umax((uint32_t)ceilf(t_max / (0.25f * H)), 2u), view_steps);
    float dt = t_max / (float)steps;
    uint32_t jitter_rng = hash_u32(rng ^ 0x3c6ef372u);
    float jitter_draw = rnd(jitter_rng);
    float jitter = (steps < view_steps || limit < exit) ? jitter_draw : 0.5f;
    const bool split = two_heights(sc);
    const float inv_hm = split ? 1.0f / sc.mie_height : 0.0f;

    uint32_t reflected_rng = hash_u32(rng ^ 0xa511e9b3u);
    ReflectorSample reflector = sample_reflector(sc, origin, reflected_rng);
    f3 reflected = f3s(0.0f), reflected_direction = dir;
    float reflected_cos = 0.0f, reflected_rayleigh = 0.0f, reflected_mie = 0.0f;
    f3 reflector_dalbedo = f3s(0.0f);  // d log(albedo) / d material, per component
    if (reflector.index != SG_MISS) {
        Ray reflected_ray = camera_ray(origin, dir);
        LightSample light = sample_light(sc, reflected_ray, reflector.index, reflected_rng);
        reflected_ray.direction = light.direction;
        reflected = reflected_radiance(sc, reflected_ray, reflector.index, reflected_rng) *
                    (reflector.inverse_probability / light.pdf);
        reflected_direction = light.direction;
        float cos_theta = dot(dir, light.direction);
        reflected_cos = cos_theta;
        reflected_rayleigh = rayleigh_phase(cos_theta);
        reflected_mie = mie_phase(cos_theta, g);
        if constexpr (G::on) {
            Hit rh = sphere_hit(sc, reflected_ray, reflector.index);
            if (rh.index != SG_MISS) {
                const Body &rb = body(sc, reflector.index);
                f3 d;
                const f3 a = surface_albedo_with_derivative(rb, rh.normal, d);
                reflector_dalbedo = f3{a.x > 0.0f ? d.x / a.x : 0.0f, a.y > 0.0f ? d.y / a.y : 0.0f,
                                       a.z > 0.0f ? d.z / a.z : 0.0f};
            }
        }
    }
    if (reflected.x > 0.0f || reflected.y > 0.0f || reflected.z > 0.0f) {
        float sun = brightest_risen_light(sc, origin);
        if (sun > 0.0f) {
            float ratio = maxc(reflected) / sun;
            float minimum = 1.0f / sqrtf(clampf((float)accumulated + 1.0f, 1.0f, 4096.0f));
            float keep = clampf(ratio * 1.0e3f, minimum, 1.0f);
            reflected = rnd(reflected_rng) < keep ? reflected / keep : f3s(0.0f);
        }
    }
    bool planetshine = reflected.x > 0.0f || reflected.y > 0.0f || reflected.z > 0.0f;

    f3 t_view = f3s(1.0f), inscatter = f3s(0.0f);
    f3 reflector_grad = f3s(0.0f);  // d(loss)/d(reflector material), summed over the steps
    float view_column = 0.0f;  // density column from the origin to this step
    uint64_t view_bins = 0u;   // density profile bins of the steps so far
    for (uint32_t i = 0; i < steps; ++i) {
        SG_COUNT(SG_EV_VIEW_CELL);
        f3 p = origin + dir * (((float)i + jitter) * dt);
        float altitude = fmaxf(length(p - center) - r, 0.0f);
        uint32_t bin = 0u;
        float density;
        if (sc.density_profile) {
            bin = density_bin(sc, altitude);
            density = sc.density_profile[bin];
        } else {
            density = sg_exp(-altitude / H);
        }
        float column = density * dt;
        f3 tau, step_t, segment;
        // Scattering coefficients of the step are (beta_r * scale_r, beta_m
        // * scale_m) times the segment. With one scale height the segment
        // is the reference's column-weighted integral and the scales are 1;
        // with two it is the transmittance integral over the step's length
        // and the scales are the two densities.
        float scale_r = 1.0f, scale_m = 1.0f;
        if (split) {
            const float density_m = sg_exp(-altitude * inv_hm);
            tau = beta_r * column + f3s(beta_m * (density_m * dt));
            step_t = exp3(-tau);
            segment.x = tau.x < 1.0e-3f ? dt * (1.0f - 0.5f * tau.x) : (1.0f - step_t.x) * dt / tau.x;
            segment.y = tau.y < 1.0e-3f ? dt * (1.0f - 0.5f * tau.y) : (1.0f - step_t.y) * dt / tau.y;
            segment.z = tau.z < 1.0e-3f ? dt * (1.0f - 0.5f * tau.z) : (1.0f - step_t.z) * dt / tau.z;
            scale_r = density;
            scale_m = density_m;
        } else {
            tau = beta_ext * column;
            step_t = exp3(-tau);
            segment.x = tau.x < 1.0e-3f ? column * (1.0f - 0.5f * tau.x) : (1.0f - step_t.x) / fmaxf(beta_ext.x, 1.0e-20f);
            segment.y = tau.y < 1.0e-3f ? column * (1.0f - 0.5f * tau.y) : (1.0f - step_t.y) / fmaxf(beta_ext.y, 1.0e-20f);
            segment.z = tau.z < 1.0e-3f ? column * (1.0f - 0.5f * tau.z) : (1.0f - step_t.z) / fmaxf(beta_ext.z, 1.0e-20f);
        }
        const f3 beta_r_step = beta_r * scale_r;
        const float beta_m_step = beta_m * scale_m;
        f3 step_dlog = f3s(0.0f);  // d log(t_view * segment) / d(beta_ext)
        f3 seg_dcol = f3s(0.0f);   // d log(segment) / d(column)
        if constexpr (G::on) {
            step_dlog = f3{segment_dlog(tau.x, column, segment.x, beta_ext.x),
                           segment_dlog(tau.y, column, segment.y, beta_ext.y),
                           segment_dlog(tau.z, column, segment.z, beta_ext.z)} -
                        f3s(view_column);
            if (sc.density_profile) {
                seg_dcol = f3{segment.x > 0.0f ? (tau.x < 1.0e-3f ? 1.0f - tau.x : step_t.x) / segment.x : 0.0f,
                              segment.y > 0.0f ? (tau.y < 1.0e-3f ? 1.0f - tau.y : step_t.y) / segment.y : 0.0f,
                              segment.z > 0.0f ? (tau.z < 1.0e-3f ? 1.0f - tau.z : step_t.z) / segment.z : 0.0f};
            }
        }
        // d(loss)/d(profile) of a weighted term of this step: the view path
        // before the step, the step's own segment, and the light column.
        auto density_terms = [&](f3 wc, const ColumnBins &light_bins) {
            if constexpr (G::on) {
                if (!sc.density_profile) return;
                const float e = -(beta_ext.x * wc.x + beta_ext.y * wc.y + beta_ext.z * wc.z);
                for (uint32_t j = 0; j < umin(i, COLUMN_BIN_SLOTS); ++j)
                    grad.density((uint32_t)((view_bins >> (4u * j)) & 15u), e * dt);
                grad.density(bin, (wc.x * seg_dcol.x + wc.y * seg_dcol.y + wc.z * seg_dcol.z) * dt);
                grad.column(sc, light_bins, e);
            }
        };
        for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
            uint32_t b = light_index(sc, slot);
            if (body(sc, b).material.w < 0.5f) continue;
            Ray light_ray = camera_ray(p, dir);
            uint32_t light_samples = eclipse_light_samples(sc, light_ray, b, accumulated);
            for (uint32_t ls = 0; ls < light_samples; ++ls) {
                LightSample light = sample_light_stratum(sc, light_ray, b, ls, light_samples, rng);
                SG_COUNT(SG_EV_LIGHT_SAMPLE);
                light_ray.direction = light.direction;
                uint32_t owed = 0u;
                float light_limit = 0.0f;
                float visibility;
                if constexpr (D::on) {
                    visibility = air_light_visibility_split(sc, light_ray, b, owed, light_limit);
                } else {
                    visibility = air_light_visibility(sc, light_ray, b);
                }
                if (visibility == 0.0f) continue;
                f3 e_sun = xyz(body(sc, b).material) * (visibility / ((float)light_samples * light.pdf));
                ColumnBins sun_bins;
                float sun_mie = 0.0f;
                float sun_column = density_integral(sc, p, light.direction, sc.air_sun_steps,
                                                   G::on ? &sun_bins : nullptr, &sun_mie);
                f3 sun_t = exp3(-air_depth(sc, sun_column, sun_mie));
                float cos_theta = dot(dir, light.direction);
                f3 beta_phase = beta_r_step * rayleigh_phase(cos_theta) + f3s(beta_m_step * mie_phase(cos_theta, g));
                f3 term = t_view * segment * beta_phase * e_sun * sun_t;
                if constexpr (D::on) {
                    if (owed) {
                        defer->push(p, light.direction, light_limit, term, owed);
                        continue;
                    }
                }
                inscatter += term;
                if constexpr (G::on) {
                    f3 wc = wt * term;
                    grad.extinction(wc * (step_dlog - f3s(sun_column)));
                    grad.phase(wc, beta_phase, rayleigh_phase(cos_theta), mie_phase(cos_theta, g),
                               mie_phase_dg(cos_theta, g), beta_m);
                    grad.shadow(sc, light_ray, b, wc.x + wc.y + wc.z);
                    density_terms(wc, sun_bins);
                }
            }
        }
        if (planetshine) {
            Ray shine_ray = camera_ray(p, reflected_direction);
            uint32_t owed = 0u;
            float shine_limit = 0.0f;
            float visibility;
            if constexpr (D::on) {
                visibility = air_light_visibility_split(sc, shine_ray, reflector.index, owed, shine_limit);
            } else {
                visibility = air_light_visibility(sc, shine_ray, reflector.index);
            }
            if (visibility > 0.0f) {
                ColumnBins light_bins;
                float light_mie = 0.0f;
                float light_column = density_integral(sc, p, reflected_direction, sc.air_sun_steps,
                                                      G::on ? &light_bins : nullptr, &light_mie);
                f3 light_t = exp3(-air_depth(sc, light_column, light_mie));
                const f3 reflected_phase = beta_r_step * reflected_rayleigh + f3s(beta_m_step * reflected_mie);
                f3 term = t_view * segment * reflected_phase * reflected * (light_t * visibility);
                bool queued = false;
                if constexpr (D::on) {
                    if (owed) {
                        defer->push(p, reflected_direction, shine_limit, term, owed);
                        queued = true;
                    }
                }
                if (!queued) inscatter += term;
                if constexpr (G::on) {
                    f3 wc = wt * term;
                    grad.extinction(wc * (step_dlog - f3s(light_column)));
                    grad.phase(wc, reflected_phase, rayleigh_phase(reflected_cos), mie_phase(reflected_cos, g),
                               mie_phase_dg(reflected_cos, g), beta_m);
                    reflector_grad += wc * reflector_dalbedo;
                    grad.shadow(sc, shine_ray, reflector.index, wc.x + wc.y + wc.z);
                    density_terms(wc, light_bins);
                }
            }
        }
        t_view *= step_t;
        if constexpr (G::on) {
            view_column += column;
            if (i < COLUMN_BIN_SLOTS) view_bins |= (uint64_t)bin << (4u * i);
        }
    }
    if constexpr (G::on) {
        if (planetshine) grad.albedo(reflector.index, reflector_grad);
    }
    result.transmittance = t_view;
    result.inscatter = inscatter;
    return result;
}
// src/shaders/pathtrace.wgsl:1217-1234
SG_FN f3 camera_air_transmittance(const Scene &sc, f3 origin, f3 dir, float limit, uint32_t seed,
                                   ColumnBins *trace = nullptr) {
    float exit = shell_exit(sc, origin, dir);
    float t_max = fminf(limit, exit);
    if (t_max <= 0.0f) return f3s(1.0f);
    float H = atmo_scale_height(sc);
    const uint32_t view_steps = sc.air_view_steps;
    uint32_t steps = umin(umax((uint32_t)ceilf(t_max / (0.25f * H)), 2u), view_steps);
    float dt = t_max / (float)steps;
    uint32_t jitter_rng = hash_u32(seed ^ 0x3c6ef372u);
    float jitter_draw = rnd(jitter_rng);
    float jitter = (steps < view_steps || limit < exit) ? jitter_draw : 0.5f;
    const bool split = two_heights(sc);
    const float inv_hm = split ? 1.0f / sc.mie_height : 0.0f;
    f3 center = atmo_center(sc);
    float r = host_radius(sc);
    f3 transmittance = f3s(1.0f);
    float total = 0.0f;
    uint64_t bins = 0u;
    for (uint32_t i = 0; i < steps; ++i) {
        f3 p = origin + dir * (((float)i + jitter) * dt);
        float altitude = fmaxf(length(p - center) - r, 0.0f);
        float column, column_m = 0.0f;
        if (sc.density_profile) {
            const uint32_t bin = density_bin(sc, altitude);
            column = sc.density_profile[bin] * dt;
            if (i < COLUMN_BIN_SLOTS) bins |= (uint64_t)bin << (4u * i);
        } else {
            column = sg_exp(-altitude / H) * dt;
            if (split) column_m = sg_exp(-altitude * inv_hm) * dt;
        }
        transmittance *= exp3(-air_depth(sc, column, column_m));
        total += column;
    }
    if (trace) {
        trace->total = total;
        trace->dt = dt;
        trace->bins = bins;
        trace->count = steps;
    }
    return transmittance;
}

// --- surface sky cache (src/shaders/forest.wgsl:3-54) ------------------------
SG_FN f3 surface_sky_texel(const Scene &sc, int x, int y) {
    float4 t = sc.sky_light[(uint32_t)imin(imax(y, 0), 7) * 16u + (uint32_t)((x + 16) % 16)];
    return f3{t.x, t.y, t.z};
}
SG_FN f3 surface_sky_lookup(const Scene &sc, f3 dir) {
    f3 p = to_ground(sc, dir);
    if (p.y <= 0.0f) return f3s(0.0f);
    float u = fmodf(atan2f(p.z, p.x) / TAU + 1.0f, 1.0f) * 16.0f - 0.5f;
    float v = p.y * 8.0f - 0.5f;
    int cx = (int)floorf(u), cy = (int)floorf(v);
    float fx = fractf(u), fy = fractf(v);
    return mix3(mix3(surface_sky_texel(sc, cx, cy), surface_sky_texel(sc, cx + 1, cy), fx),
                mix3(surface_sky_texel(sc, cx, cy + 1), surface_sky_texel(sc, cx + 1, cy + 1), fx), fy);
}
SG_FN bool sky_blocked(const Scene &sc, const Ray &ray) {
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (sc.s.globals.ground_north[3] == (float)(i + 1u)) continue;
        if (sphere_hit(sc, ray, i).index != SG_MISS) return true;
    }
    return landscape_query(sc, ray, 1.0e30f, true).index != SG_MISS;
}
SG_FN f3 surface_sky(const Scene &sc, const Ray &origin, f3 normal, uint32_t &rng) {
    float u = rnd(rng);
    float phi = TAU * rnd(rng);
    f3 dir = basis_direction(normal, sqrtf(u) * cosf(phi), sqrtf(u) * sinf(phi), sqrtf(1.0f - u));
    f3 light = surface_sky_lookup(sc, dir);
    if (maxc(light) <= 0.0f) return f3s(0.0f);
    Ray ray{origin.offset, dir, origin.anchor};
    if (sky_blocked(sc, ray)) return f3s(0.0f);
    float transmission = 1.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        RingHit hit = ring_hit(sc, ray, i);
        if (hit.index != SG_MISS) transmission *= ring_transmission(sc, ray, hit);
    }
    return light * transmission;
}
// One texel of the 16x8 sky cache (src/shaders/forest.wgsl:3-17).
SG_FN f3 sky_cache_texel(const Scene &sc, uint32_t x, uint32_t y) {
    f3 light = f3s(0.0f);
    if (sc.s.globals.atmo_params[3] > 0.5f && sc.s.globals.ground_counts[1] != 0u) {
        uint32_t index = y * 16u + x;
        float mu = ((float)y + 0.5f) / 8.0f;
        float phi = TAU * ((float)x + 0.5f) / 16.0f;
        float r = sqrtf(1.0f - mu * mu);
        f3 dir = from_ground(sc, f3{r * cosf(phi), mu, r * sinf(phi)});
        uint32_t rng = hash_u32(index ^ 0x91ce452du);
        light = atmosphere(sc, f3s(0.0f), dir, 1.0e30f, rng, sc.s.counts[1]).inscatter;
    }
    return light;
}

// --- ringshine / point discs (src/shaders/pathtrace.wgsl:1237-1345) ---------
SG_FN float ring_area(const Scene &sc, uint32_t index) {
    const Body &b = body(sc, index);
    float r = b.center.w;
    return PI * r * r * (b.ring_params.x * b.ring_params.x - b.ring_plane.w * b.ring_plane.w);
}
// With a gradient sink, `wt` is d(loss)/d(ringshine) and the derivatives
// through the tabulated ring profile (at the sampled ring point and on its
// shadow rays) are added to `grad`.
template <class G = NoGrad>
SG_FN f3 ringshine(const Scene &sc, uint32_t index, f3 origin, f3 normal, bool last_vertex, uint32_t &rng,
                   const G &grad = G(), f3 wt = f3s(0.0f)) {
    const Body &b = body(sc, index);
    if (b.ring_params.x <= b.ring_plane.w || b.ring_params.y <= 0.0f) return f3s(0.0f);
    f3 n = xyz(b.ring_plane);
    float inner = b.ring_plane.w * b.ring_plane.w;
    float outer = b.ring_params.x * b.ring_params.x;
    float radius = sqrtf(mixf(inner, outer, rnd(rng))) * b.center.w;
    float phi = TAU * rnd(rng);
    f3 point = basis_direction(n, cosf(phi), sinf(phi), 0.0f) * radius;
    f3 offset = point - origin;
    float distance2 = dot(offset, offset);
    float distance = sqrtf(distance2);
    f3 direction = offset / distance;
    float cos_surface = dot(normal, direction);
    float cos_ring = fabsf(dot(direction, n));
    if (cos_surface <= 0.0f || cos_ring < 1.0e-6f) return f3s(0.0f);
    Hit blocker = sphere_hit(sc, Ray{tod(origin), direction, index}, index);
    if (blocker.index != SG_MISS && blocker.distance < distance) return f3s(0.0f);
    float4 optics = ring_optics(sc, index, point);
    if (optics.w == 0.0f) return f3s(0.0f);
    float mu_out = dot(-direction, n);
    float pdf = distance2 / (ring_area(sc, index) * cos_ring);
    float weight = last_vertex ? 1.0f : power_weight(pdf, cos_surface / PI);
    const float scale = cos_surface * weight / (PI * pdf);
    f3 shine = f3s(0.0f);
    const bool tabulated = G::on && index == sc.ring_profile_body;
    f3 shine_tau = f3s(0.0f);  // d shine / d tau at the sampled point
    for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
        uint32_t i = light_index(sc, slot);
        if (body(sc, i).material.w < 0.5f) continue;
        LightSample light = sample_light(sc, Ray{tod(point), n, index}, i, rng);
        float mu_in = dot(light.direction, n);
        Ray shadow = ring_origin(sc, RingHit{0.0f, index, point}, light.direction);
        float visibility = light_visibility(sc, shadow, i);
        f3 term = ring_bsdf(optics, mu_out, mu_in) * xyz(body(sc, i).material) * (fabsf(mu_in) * visibility / light.pdf);
        shine += term;
        if constexpr (G::on) {
            if (tabulated) {
                float2 f = ring_bsdf_factor(optics.w, mu_out, mu_in);
                shine_tau += xyz(optics) * xyz(body(sc, i).material) *
                             (f.y / (4.0f * PI) * fabsf(mu_in) * visibility / light.pdf);
            }
            f3 wc = wt * term * scale;
            grad.shadow(sc, shadow, i, wc.x + wc.y + wc.z);
        }
    }
    if constexpr (G::on) {
        if (tabulated) {
            f3 wc = wt * shine_tau * scale;
            grad.ring(ring_profile_bin(ring_fraction(b, point)), wc.x + wc.y + wc.z);
        }
    }
    return shine * scale;
}
SG_FN float pixel_angle(const Scene &sc) {
    return 2.0f * sc.s.globals.cam_forward[3] / fmaxf(sc.s.globals.viewport[1], 1.0f);
}
SG_FN Hit point_disc_hit(const Scene &sc, const Ray &ray, const Hit &closest, float &scale, uint32_t &rng) {
    Hit result = closest;
    float total = 0.0f;
    float pixel = pixel_angle(sc);
    float angle = fminf(fmaxf(POINT_DISC_RADIUS, 0.7f * pixel), STAR_DOT_PIXELS * pixel);
    uint32_t host = (uint32_t)sc.s.globals.atmo_params[3] - 1u;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (i == host || sc.s.globals.ground_north[3] == (float)(i + 1u)) continue;
        float radius = body(sc, i).center.w;
        float drawn = angle * length(tof(relative_center(sc, ray, i)));
        if (radius <= 0.0f || drawn <= radius) continue;
        Hit hit = sphere_hit_radius(sc, ray, i, drawn);
        if (hit.distance < closest.distance) {
            float ratio = (radius / drawn) * (radius / drawn);
            total += ratio;
            if (rnd(rng) * total < ratio) result = hit;
        }
    }
    if (total > 0.0f) scale = total;
    return result;
}
SG_FN float point_gain(const Scene &sc, float extent, float distance) {
    float pixels = 2.0f * extent / fmaxf(distance, 1.0e-30f) / pixel_angle(sc);
    return mixf(sc.s.background[3], 1.0f, smoothstepf(1.0f, 3.0f, pixels));
}

namespace {
// --- the integrator (src/shaders/pathtrace.wgsl:1347-1617) -----------------
// trace() is split at its loop boundaries into steps that carry the whole
// path in PathState:
//   path_begin       initial state for one camera sample
//   path_extend      closest hit, depth-0 point discs (consumes draws), closest ring
//   path_camera_air  first segment only: the camera-air pass folded in
//                    (integrate_camera_air, :1630-1643, traced the same camera
//                    ray with the same RNG state, so its in-scatter and its
//                    continuation RNG are computed here instead of read back)
//   path_shade       the rest of one loop iteration; false when the path ends
//   path_finish      ring indirect cap and the returned vec4
// The fused kernel loops the steps per lane; the wavefront kernels run each
// step over a compacted queue. Both consume draws in the reference's order.
enum : uint32_t {
    PATH_CAMERA_SEGMENT = 1u << 0,
    PATH_IN_AIR = 1u << 1,
    PATH_SKY_ON_ESCAPE = 1u << 2,
    PATH_POINT_GAINED = 1u << 3,
};
struct PathState {
    Ray ray;
    Ray previous_vertex;
    f3 throughput, radiance, ring_base, backdrop;
    float previous_pdf, ring_cap;
    uint32_t rng, depth, null_crossings, shine_body, flags, accumulated;
};
// Path replay adjoint state: the primal run's returned radiance and the loss
// weight w = d(loss)/d(radiance) of the path sample.
//
// The ring indirect cap: where a path first scatters off a ring the cap
// c = max(direct, RING_INDIRECT_CAP * unshadowed) is set from that event's
// lights and the radiance so far becomes the base B; the returned radiance is
// B + I * min(1, c / lum(I)) for the indirect radiance I gathered after it.
// When the cap binds (s = c / lum(I) < 1) the return is B + I * c / lum(I),
// whose derivative the replay takes in full: with alpha = dot(w, I) / lum(I),
//  - terms added after the cap weigh s * (w - alpha * lum_weights), the
//    luminance of the capped part being held at c;
//  - a throughput factor before the cap scales B's rest, I and c alike, and
//    so weighs w * (value - radiance so far) + alpha * lum_weights *
//    (D - s I), D being the colour of the binding cap term;
//  - a factor after it scales only I's rest;
//  - the cap event's own light terms add alpha * d c.
// rest() is d(loss)/d log(a throughput factor applied now).
struct PathGrad : GradSink {
    f3 w;
    f3 value;
    bool capped = false;
    bool cap_via_direct = false;
    uint32_t cap_depth = 0u, cap_nulls = 0u;
    float alpha = 0.0f, scale = 1.0f;
    f3 wl = f3{0.0f, 0.0f, 0.0f};        // alpha * luminance weights
    f3 indirect = f3{0.0f, 0.0f, 0.0f};  // s * I, the capped indirect radiance
    f3 before = f3{0.0f, 0.0f, 0.0f};    // alpha * lum_weights * (D - s I)
    SG_MEMBER bool past_cap(const PathState &st) const { return capped && st.ring_cap < 1.0e29f; }
    SG_MEMBER f3 weight(const PathState &st) const { return past_cap(st) ? (w - wl) * scale : w; }
    SG_MEMBER f3 rest(const PathState &st) const {
        if (!capped) return w * (value - st.radiance);
        if (st.ring_cap >= 1.0e29f) return w * (value - st.radiance) + before;
        return (w - wl) * (indirect - (st.radiance - st.ring_base) * scale);
    }
    // Whether the ring event the path is at sets the cap that bound.
    SG_MEMBER bool cap_event(const PathState &st) const {
        return capped && st.ring_cap >= 1.0e29f && st.depth == cap_depth && st.null_crossings == cap_nulls;
    }
    // The cap's make-up from the primal run `st` that `cap` traced.
    SG_MEMBER void set_cap(const PathState &st, const CapTrace &cap) {
        capped = false;
        if (!cap.set || st.ring_cap >= 1.0e29f) return;
        const f3 in = st.radiance - st.ring_base;
        const float lum = luminance(in);
        if (!(lum > st.ring_cap)) return;
        capped = true;
        cap_via_direct = cap.via_direct;
        cap_depth = cap.depth;
        cap_nulls = cap.nulls;
        scale = st.ring_cap / lum;
        alpha = dot(w, in) / lum;
        wl = f3{0.2126f, 0.7152f, 0.0722f} * alpha;
        indirect = in * scale;
        before = wl * (cap.vector - indirect);
    }
};
}  // namespace

struct HitRecord {
    Hit hit;
    float disc_scale;
    RingHit ring;
};
struct PathResult {
    float4 value;  // rgb radiance, backdrop luminance
    f3 backdrop;
};

SG_FN void path_begin(PathState &st, f3 direction, uint32_t rng, uint32_t accumulated) {
    st.ray = Ray{d3{0.0, 0.0, 0.0}, direction, SG_MISS};
    st.previous_vertex = st.ray;
    st.throughput = f3s(1.0f);
    st.radiance = f3s(0.0f);
    st.ring_base = f3s(0.0f);
    st.backdrop = f3s(0.0f);
    st.previous_pdf = 0.0f;
    st.ring_cap = 1.0e30f;
    st.rng = rng;
    st.depth = 0u;
    st.null_crossings = 0u;
    st.shine_body = SG_MISS;
    st.flags = PATH_CAMERA_SEGMENT;
    st.accumulated = accumulated;
}

SG_FN HitRecord path_extend(const Scene &sc, PathState &st) {
    HitRecord h;
    h.hit = closest_hit(sc, st.ray);
    h.disc_scale = 1.0f;
    if (st.depth == 0u && st.null_crossings == 0u && sc.s.background[3] > 0.0f)
        h.hit = point_disc_hit(sc, st.ray, h.hit, h.disc_scale, st.rng);
    h.ring = closest_ring(sc, st.ray, h.hit.distance);
    return h;
}

template <class G = NoGrad, class D = NoDefer>
SG_FN void path_camera_air(const Scene &sc, PathState &st, const HitRecord &h, const G &grad = G(),
                           D *defer = nullptr) {
    if (!(st.flags & PATH_CAMERA_SEGMENT)) return;
    st.flags &= ~PATH_CAMERA_SEGMENT;
    if (sc.s.globals.atmo_params[3] < 0.5f) return;
    float limit = 1.0e30f;
    if (h.hit.index != SG_MISS) limit = h.hit.distance;
    if (h.ring.index != SG_MISS) limit = fminf(limit, h.ring.distance);
    f3 origin = tof(st.ray.offset);
    uint32_t air_rng = st.rng;
    Scatter air;
    ColumnBins column;
    if constexpr (G::on) {
        air = atmosphere(sc, origin, st.ray.direction, fminf(h.hit.distance, h.ring.distance), air_rng,
                         st.accumulated, static_cast<const GradSink &>(grad), grad.weight(st) * st.throughput);
    } else {
        air = atmosphere<NoGrad, D>(sc, origin, st.ray.direction, fminf(h.hit.distance, h.ring.distance), air_rng,
                                    st.accumulated, NoGrad(), f3s(0.0f), defer);
    }
    f3 transmittance = camera_air_transmittance(sc, origin, st.ray.direction, limit, st.rng,
                                                G::on ? &column : nullptr);
    st.radiance += st.throughput * air.inscatter;
    if constexpr (G::on) {
        f3 r = grad.rest(st);
        grad.extinction(r * (-column.total));
        f3 be = beta_rayleigh(sc) + f3s(beta_mie(sc));
        grad.column(sc, column, -(be.x * r.x + be.y * r.y + be.z * r.z));
    }
    st.throughput *= transmittance;
    st.rng = air_rng;
}

// With a PathGrad, replays the step of the primal run and adds the
// derivatives of every contribution and throughput factor with respect to
// body materials, the tabulated ring profile and the air coefficients.
// Discrete choices whose probability the weights compensate (Russian
// roulette, reflector and point-disc selection, ring scattering) carry no
// score; the ring pass-through, weighted 1 with probability T, carries
// d log T. MIS weights are taken as constants.
template <class G = NoGrad>
SG_FN bool path_shade(const Scene &sc, PathState &st, const HitRecord &h, const G &grad = G()) {
    const sg_settings &s = sc.s;
    const uint32_t host = (uint32_t)s.globals.atmo_params[3] - 1u;
    const uint32_t max_vertices = s.counts[3];
    const Hit &hit = h.hit;
    const RingHit &ring = h.ring;
    Ray &ray = st.ray;
    if (st.flags & PATH_IN_AIR) {
        st.flags &= ~PATH_IN_AIR;
        if (ring.index != SG_MISS || (hit.index < SG_LOCAL_BASE && hit.index != host)) {
            ColumnBins bins;
            float mie = 0.0f;
            float column = air_column(sc, ray, ray.direction, G::on ? &bins : nullptr, &mie);
            if constexpr (G::on) {
                if (column > 0.0f) {
                    f3 r = grad.rest(st);
                    grad.extinction(r * (-column));
                    f3 be = beta_rayleigh(sc) + f3s(beta_mie(sc));
                    grad.column(sc, bins, -(be.x * r.x + be.y * r.y + be.z * r.z));
                }
            }
            st.throughput *= air_extinction(sc, column, mie);
        }
    }
    if (ring.index != SG_MISS) {
        const Body &planet = body(sc, ring.index);
        if (st.depth == 0u && !(st.flags & PATH_POINT_GAINED) && s.background[3] > 0.0f) {
            st.flags |= PATH_POINT_GAINED;
            st.throughput *= point_gain(sc, planet.center.w * planet.ring_params.x, ring.distance);
        }
        f3 n = xyz(planet.ring_plane);
        float4 optics = ring_optics(sc, ring.index, ring.point);
        float mu_out = dot(-ray.direction, n);
        float transmission = ring_transmission(sc, ray, ring);
        float scatter_probability = one_minus_exp(optics.w / fmaxf(fabsf(mu_out), 1.0e-7f));
        float bsdf_pdf = scatter_probability / (4.0f * PI);
        bool last_vertex = st.depth + 1u == max_vertices;
        float ring_mis = 1.0f;
        if (st.shine_body == ring.index && st.null_crossings == 0u) {
            float cos_ring = fmaxf(fabsf(dot(ray.direction, n)), 1.0e-6f);
            float shine_pdf = ring.distance * ring.distance / (ring_area(sc, ring.index) * cos_ring);
            ring_mis = power_weight(st.previous_pdf, shine_pdf);
        }
        st.shine_body = SG_MISS;
        const bool tabulated = G::on && ring.index == sc.ring_profile_body;
        uint32_t bin = 0u;
        if constexpr (G::on) bin = ring_profile_bin(ring_fraction(planet, ring.point));
        float unshadowed = 0.0f, direct = 0.0f;
        f3 direct_colour = f3s(0.0f), unshadowed_colour = f3s(0.0f);  // for CapTrace
        bool cap_event = false;
        if constexpr (G::on) cap_event = grad.cap_event(st);
        for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
            uint32_t i = light_index(sc, slot);
            if (body(sc, i).material.w < 0.5f || optics.w == 0.0f) continue;
            LightSample light = sample_light(sc, Ray{tod(ring.point), n, ring.index}, i, st.rng);
            float mu_in = dot(light.direction, n);
            Ray shadow = ring_origin(sc, ring, light.direction);
            float visibility = light_visibility(sc, shadow, i);
            float weight = last_vertex ? 1.0f : power_weight(light.pdf, bsdf_pdf);
            f3 lit = st.throughput * ring_bsdf(optics, mu_out, mu_in) * xyz(body(sc, i).material) *
                     (fabsf(mu_in) / light.pdf);
            if constexpr (G::on) {
                f3 wc = grad.weight(st) * lit * (visibility * weight * ring_mis);
                grad.shadow(sc, shadow, i, wc.x + wc.y + wc.z);
                if (tabulated) {
                    float2 f = ring_bsdf_factor(optics.w, mu_out, mu_in);
                    f3 d = grad.weight(st) * st.throughput * xyz(optics) * xyz(body(sc, i).material) *
                           (f.y / (4.0f * PI) * fabsf(mu_in) / light.pdf * visibility * weight * ring_mis);
                    grad.ring(bin, d.x + d.y + d.z);
                }
                if (cap_event) {
                    // alpha * d c: the cap's light terms (the throughput in
                    // them is the pre-cap rest's).
                    const float l = grad.alpha * luminance(lit);
                    if (grad.cap_via_direct) grad.shadow(sc, shadow, i, l * visibility);
                    if (tabulated) {
                        float2 f = ring_bsdf_factor(optics.w, mu_out, mu_in);
                        if (f.x > 0.0f)
                            grad.ring(bin, l * (f.y / f.x) * (grad.cap_via_direct ? visibility : RING_INDIRECT_CAP));
                    }
                }
            }
            if constexpr (G::capture) {
                direct_colour += lit * visibility;
                unshadowed_colour += lit;
            }
            st.radiance += lit * (visibility * weight * ring_mis);
            unshadowed += luminance(lit);
            direct += luminance(lit) * visibility;
        }
        if (rnd(st.rng) < transmission) {
            if constexpr (G::on) {
                if (tabulated) {
                    f3 r = grad.rest(st);
                    grad.ring(bin, (r.x + r.y + r.z) * (-1.0f / fmaxf(fabsf(dot(ray.direction, n)), 1.0e-7f)));
                }
            }
            ray = ring_origin(sc, ring, ray.direction);
            st.null_crossings += 1u;
            return st.null_crossings <= s.counts[0];
        }
        if (last_vertex || scatter_probability == 0.0f) return false;
        if (st.ring_cap > 1.0e29f) {
            st.ring_cap = fmaxf(direct, RING_INDIRECT_CAP * unshadowed);
            st.ring_base = st.radiance;
            if constexpr (G::capture) {
                grad.set = true;
                grad.via_direct = direct >= RING_INDIRECT_CAP * unshadowed;
                grad.vector = grad.via_direct ? direct_colour : unshadowed_colour * RING_INDIRECT_CAP;
                grad.depth = st.depth;
                grad.nulls = st.null_crossings;
            }
        }
        float z = 1.0f - 2.0f * rnd(st.rng);
        float phi = TAU * rnd(st.rng);
        float radial = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        f3 direction = basis_direction(n, radial * cosf(phi), radial * sinf(phi), z);
        if constexpr (G::on) {
            if (tabulated) {
                float2 f = ring_bsdf_factor(optics.w, mu_out, z);
                if (f.x > 0.0f) {
                    f3 r = grad.rest(st);
                    grad.ring(bin, (r.x + r.y + r.z) * (f.y / f.x));
                }
            }
        }
        st.throughput *= ring_bsdf(optics, mu_out, z) * (fabsf(z) / bsdf_pdf);
        st.previous_pdf = bsdf_pdf;
        st.flags &= ~PATH_SKY_ON_ESCAPE;
        ray = ring_origin(sc, ring, direction);
        st.previous_vertex = ray;
        st.depth += 1u;
        st.null_crossings = 0u;
        if (st.depth >= 3u) {
            float survival = clampf(maxc(st.throughput), 0.05f, 0.95f);
            if (rnd(st.rng) >= survival) return false;
            st.throughput /= survival;
        }
        return true;
    }
    if (hit.index == SG_MISS) {
        st.radiance += st.throughput * environment(sc, ray.direction, st.depth == 0u);
        if (st.flags & PATH_SKY_ON_ESCAPE) st.radiance += st.throughput * surface_sky_lookup(sc, ray.direction);
        if (st.depth == 0u) {
            st.backdrop = milky_way(sc, ray.direction) * st.throughput;
            st.radiance += st.backdrop;
        }
        return false;
    }
    const bool local = kLandscape && hit.index >= SG_LOCAL_BASE;
    st.throughput *= h.disc_scale;
    const Body *hb = local ? nullptr : &sc.bodies[hit.index];
    if (!local && hb->material.w > 0.5f) {
        float weight = 1.0f;
        if (st.depth > 0u) {
            float light_pdf = 1.0f / (TAU * cone_width(sc, st.previous_vertex, hit.index));
            weight = power_weight(st.previous_pdf, light_pdf);
        }
        st.radiance += st.throughput * xyz(hb->material) * weight;
        return false;
    }
    if (!local && st.depth == 0u && !(st.flags & PATH_POINT_GAINED) && s.background[3] > 0.0f) {
        st.flags |= PATH_POINT_GAINED;
        float extent = hb->center.w * fmaxf(1.0f, hb->ring_params.x);
        st.throughput *= point_gain(sc, extent, hit.distance);
    }
    f3 normal = local ? landscape_shading_normal(sc, ray, hit) : hit.normal;
    Ray outgoing = local ? landscape_outgoing(sc, ray, hit)
                         : Ray{tod(normal * (hb->center.w * (1.0f + 2.0e-6f))), normal, hit.index};
    // Material derivative and the weight of this vertex's direct light.
    f3 dalbedo = f3s(0.0f), wt_direct = f3s(0.0f), body_albedo = f3s(0.0f);
    if constexpr (G::on) {
        if (!local) body_albedo = surface_albedo_with_derivative(*hb, normal, dalbedo);
        wt_direct = grad.weight(st) * st.throughput * body_albedo;
    }
    f3 direct_light = f3s(0.0f);
    bool last_vertex = st.depth + 1u == max_vertices;
    for (uint32_t slot = 0; slot < light_count(sc); ++slot) {
        uint32_t i = light_index(sc, slot);
        if (body(sc, i).material.w < 0.5f) continue;
        LightSample sample = sample_light(sc, outgoing, i, st.rng);
        float cosine = fmaxf(dot(normal, sample.direction), 0.0f);
        if (cosine <= 0.0f) continue;
        outgoing.direction = sample.direction;
        float visibility = light_visibility(sc, outgoing, i);
        if (visibility > 0.0f) {
            float bsdf_pdf = cosine / PI;
            float weight = last_vertex ? 1.0f : power_weight(sample.pdf, bsdf_pdf);
            ColumnBins bins;
            float mie = 0.0f;
            float column = air_column(sc, outgoing, sample.direction, G::on ? &bins : nullptr, &mie);
            f3 term = xyz(body(sc, i).material) * air_extinction(sc, column, mie) *
                      (bsdf_pdf * weight * visibility / sample.pdf);
            direct_light += term;
            if constexpr (G::on) {
                f3 wc = wt_direct * term;
                if (column > 0.0f) {
                    grad.extinction(wc * (-column));
                    f3 be = beta_rayleigh(sc) + f3s(beta_mie(sc));
                    grad.column(sc, bins, -(be.x * wc.x + be.y * wc.y + be.z * wc.z));
                }
                grad.shadow(sc, outgoing, i, wc.x + wc.y + wc.z);
            }
        }
    }
    f3 albedo;
    if constexpr (G::on) albedo = local ? landscape_albedo(sc, ray, hit) : body_albedo;
    else albedo = local ? landscape_albedo(sc, ray, hit) : surface_albedo(*hb, normal);
    if constexpr (G::on) {
        if (!local) grad.albedo(hit.index, grad.weight(st) * st.throughput * dalbedo * direct_light);
    }
    st.radiance += st.throughput * albedo * direct_light;
    st.shine_body = SG_MISS;
    if (!local) {
        f3 shine;
        if constexpr (G::on) {
            shine = ringshine(sc, hit.index, tof(outgoing.offset), normal, last_vertex, st.rng,
                              static_cast<const GradSink &>(grad), wt_direct);
            grad.albedo(hit.index, grad.weight(st) * st.throughput * dalbedo * shine);
        } else {
            shine = ringshine(sc, hit.index, tof(outgoing.offset), normal, last_vertex, st.rng);
        }
        st.radiance += st.throughput * albedo * shine;
        st.shine_body = hit.index;
    }
    bool sky_lit = local && s.globals.ground_counts[1] != 0u && s.globals.atmo_params[3] > 0.5f;
    if (sky_lit && last_vertex) st.radiance += st.throughput * albedo * surface_sky(sc, outgoing, normal, st.rng);
    if (last_vertex) return false;
    float u = rnd(st.rng);
    float phi = TAU * rnd(st.rng);
    float cosine = sqrtf(1.0f - u);
    outgoing.direction = basis_direction(normal, sqrtf(u) * cosf(phi), sqrtf(u) * sinf(phi), cosine);
    st.previous_pdf = cosine / PI;
    if constexpr (G::on) {
        if (!local) {
            f3 r = grad.rest(st);
            grad.albedo(hit.index, f3{albedo.x > 0.0f ? r.x * dalbedo.x / albedo.x : 0.0f,
                                      albedo.y > 0.0f ? r.y * dalbedo.y / albedo.y : 0.0f,
                                      albedo.z > 0.0f ? r.z * dalbedo.z / albedo.z : 0.0f});
        }
    }
    st.throughput *= albedo;
    if (st.depth >= 2u) {
        float survival = clampf(maxc(st.throughput), 0.05f, 0.95f);
        if (rnd(st.rng) >= survival) return false;
        st.throughput /= survival;
    }
    ray = outgoing;
    st.previous_vertex = outgoing;
    if (inside_air(sc, outgoing)) st.flags |= PATH_IN_AIR;
    if (sky_lit) st.flags |= PATH_SKY_ON_ESCAPE;
    else st.flags &= ~PATH_SKY_ON_ESCAPE;
    st.depth += 1u;
    st.null_crossings = 0u;
    return true;
}

SG_FN PathResult path_finish(const PathState &st) {
    f3 radiance = st.radiance;
    if (st.ring_cap < 1.0e29f) {
        f3 indirect = radiance - st.ring_base;
        float lum = luminance(indirect);
        if (lum > st.ring_cap) radiance = st.ring_base + indirect * (st.ring_cap / lum);
    }
    PathResult out;
    out.value = make_float4(radiance.x, radiance.y, radiance.z, luminance(st.backdrop));
    out.backdrop = st.backdrop;
    return out;
}

// The loop of trace() over one path (its state already begun).
template <class G = NoGrad>
SG_FN void run_path(const Scene &sc, PathState &st, const G &grad = G()) {
    for (;;) {
        HitRecord h = path_extend(sc, st);
        path_camera_air(sc, st, h, grad);
        if (!path_shade(sc, st, h, grad)) break;
    }
}
// Whole path in one lane.
SG_FN PathResult trace_path(const Scene &sc, f3 direction, uint32_t rng, uint32_t accumulated) {
    PathState st;
    path_begin(st, direction, rng, accumulated);
    run_path(sc, st);
    return path_finish(st);
}
// Scale path_finish applies to the indirect radiance after the ring cap.
SG_FN float ring_cap_scale(const PathState &st) {
    if (st.ring_cap >= 1.0e29f) return 1.0f;
    float lum = luminance(st.radiance - st.ring_base);
    return lum > st.ring_cap ? st.ring_cap / lum : 1.0f;
}
// The replay half of a path sample's adjoint: given the primal run's returned
// radiance and ring-cap scale, rerun the path with the same draws adding
// d(loss)/d(parameters) for d(loss)/d(radiance) = w. A path whose cap bound
// (scale below 1) first reruns its primal to learn the cap's make-up.
SG_FN void replay_path(const Scene &sc, f3 direction, uint32_t rng, uint32_t accumulated, f3 w, f3 value,
                       float cap_scale, float *gradient, float *density, uint32_t density_stride) {
    PathState st;
    PathGrad grad;
    grad.g = gradient;
    grad.dens = density;
    grad.dstride = density_stride;
    grad.w = w;
    grad.value = value;
    if (cap_scale < 1.0f) {
        CapTrace cap;
        path_begin(st, direction, rng, accumulated);
        run_path(sc, st, cap);
        grad.set_cap(st, cap);
    }
    path_begin(st, direction, rng, accumulated);
    run_path(sc, st, grad);
    grad.flush_private();
}
// One path sample's adjoint: the primal run, then its replay.
// `density` holds GRAD_PRIVATE zeroed lane-private floats, `density_stride`
// apart, for the density-profile and air-coefficient terms; they are moved
// into `gradient` (and zeroed) at the end.
SG_FN PathResult adjoint_path(const Scene &sc, f3 direction, uint32_t rng, uint32_t accumulated, f3 w,
                              float *gradient, float *density, uint32_t density_stride) {
    PathState st;
    CapTrace cap;
    path_begin(st, direction, rng, accumulated);
    run_path(sc, st, cap);
    PathResult primal = path_finish(st);
    PathGrad grad;
    grad.g = gradient;
    grad.dens = density;
    grad.dstride = density_stride;
    grad.w = w;
    grad.value = f3{primal.value.x, primal.value.y, primal.value.z};
    grad.set_cap(st, cap);
    path_begin(st, direction, rng, accumulated);
    run_path(sc, st, grad);
    grad.flush_private();
    return primal;
}

// src/shaders/pathtrace.wgsl:1619-1626
SG_FN f3 camera_direction(const Scene &sc, uint32_t x, uint32_t y, uint32_t &rng) {
    float jx = rnd(rng);
    float jy = rnd(rng);
    float u = ((float)x + jx) / sc.s.globals.viewport[0];
    float v = ((float)y + jy) / sc.s.globals.viewport[1];
    float ndc_x = 2.0f * u - 1.0f, ndc_y = 1.0f - 2.0f * v;
    float tangent = sc.s.globals.cam_forward[3];
    return normalize(f3{ndc_x * tangent * sc.s.globals.viewport[3], ndc_y * tangent, -1.0f});
}
// Path seed for (pixel, sample) (src/shaders/pathtrace.wgsl:1635,1656).
SG_FN uint32_t path_seed(uint32_t pixel, uint32_t sequence) { return hash_u32(pixel ^ hash_u32(sequence + 0x9e3779b9u)); }

}  // namespace sg
