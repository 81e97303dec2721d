// Device-side restatement of stargaze's path-tracing helpers for gfx950.
//
// Each function cites the WGSL it stands in for (paths relative to the
// reference repository). Differences that are deliberate:
//  * Body centres are rebased in FP64 from their f32 high/low parts, and ray
//    origins are kept in FP64 (`Ray::offset`), so the camera-relative
//    rebasing of src/shaders/pathtrace.wgsl:100-113 loses nothing.
//  * Sphere quadratics for a ray that starts within four radii of the centre
//    (the observer's host planet, bounce rays leaving a body) run in FP64;
//    distant bodies keep the reference's f32 cross-product form.
//  * The metre-frame host-sphere quadratic of src/shaders/ground.wgsl:138-158
//    and the 5 mm outgoing offset of :649-662 run in FP64.
// Everything else is FP32, as in the reference.
#pragma once
#include "sg_math.h"
#include "../../include/stargaze_amd.h"

#define SG_MISS 0xffffffffu
#define SG_LOCAL_BASE 0x80000000u   // src/shaders/ground.wgsl:14
#define SG_TERRAIN_ID 0x80100000u   // src/shaders/terrain.wgsl:23

// Per-lane traversal stacks (scratch memory). The terrain stack holds
// 3 * levels + 1 entries: 12 internal levels (8192 cells) need 37.
#define SG_TERRAIN_STACK 40
#define SG_BVH_STACK 48  // the host builder's median split stays far shallower

namespace sg {

constexpr float PI = 3.141592653589793f;
constexpr float TAU = 6.283185307179586f;
constexpr float METRES_PER_AU = 149597870700.0f;   // src/shaders/ground.wgsl:15
constexpr double METRES_PER_AU_D = 149597870700.0;
constexpr float STAR_DOT_PIXELS = 3.0f;            // src/shaders/pathtrace.wgsl:65
// Whether this translation unit renders the landscape (see SG_LANDSCAPE).
constexpr bool kLandscape = SG_LANDSCAPE != 0;
// Default march of the single-scattering air: view cells per camera ray and
// density samples per light path (src/shaders/pathtrace.wgsl:67-68). A frame
// may choose others (Scene::air_view_steps / air_sun_steps), as the
// reference's own shader variants do (src/render_experiments.rs:48-70).
constexpr uint32_t ATMO_VIEW_STEPS = 8u;
constexpr uint32_t ATMO_SUN_STEPS = 4u;
constexpr float POINT_DISC_RADIUS = 0.00136f;      // src/shaders/pathtrace.wgsl:1300
constexpr float RING_INDIRECT_CAP = 0.003f;        // src/shaders/pathtrace.wgsl:1336

// WGSL `Body` (src/shaders/pathtrace.wgsl:26-37), byte-identical to sg_sphere.
struct Body {
    float4 center, material, low, ring_plane, ring_params, ring_color, axis_x, axis_y,
        surface_color, surface_params;
};
static_assert(sizeof(Body) == 160, "Body layout");
static_assert(sizeof(sg_sphere) == 160, "Sphere layout");
static_assert(sizeof(sg_globals) == 176, "Globals layout");
static_assert(sizeof(sg_settings) == 240, "Settings layout");
static_assert(sizeof(sg_primitive) == 48, "Primitive layout");
static_assert(sizeof(sg_ray_star) == 32, "RayStar layout");
static_assert(sizeof(sg_camera_air) == 16, "CameraAir layout");

// Terrain buffer geometry derived from ground_counts.z (src/shaders/terrain.wgsl:6-20).
// `data` floats come from the caller in the reference's layout (heights,
// derivatives when cells <= 1024, min/max per node). The caches after them
// are this library's own: node AABBs padded to 8 floats (two aligned
// float4 loads per node) and vertices with the reference's stride.
struct TerrainLayout {
    uint32_t cells, width, compact, stride;
    uint64_t vertices, nodes, bounds, data, cache, vertex_cache, total;
};

SG_FN TerrainLayout terrain_layout(uint32_t ground_counts_z) {
    TerrainLayout t;
    uint32_t log2 = ground_counts_z >> 8;
    t.cells = log2 != 0 ? (1u << log2) : 512u;
    t.width = t.cells + 1;
    t.compact = t.cells > 1024u;
    t.stride = t.compact ? 2u : 6u;
    t.vertices = (uint64_t)t.width * t.width;
    uint64_t leaves = t.cells / 2u;
    t.nodes = (4u * leaves * leaves - 1u) / 3u;
    t.bounds = (t.compact ? 1u : 3u) * t.vertices;
    t.data = t.bounds + 2u * t.nodes;
    t.cache = (t.data + 3u) & ~(uint64_t)3u;
    t.vertex_cache = t.cache + 8u * t.nodes;
    t.total = t.vertex_cache + (uint64_t)t.stride * t.vertices;
    return t;
}

// Everything a kernel reads about the scene. Passed by value as a kernel
// argument, so per-scene constants come through scalar loads.
// A body's centre in FP64 (high + low parts of its f32 record), 32 B.
struct Centre {
    double x, y, z, pad;
};
struct Scene {
    sg_settings s;
    const Body *bodies;
    // FP64 centres of `bodies`, formed once per frame upload (fill_centres)
    // instead of in every lane of every body test.
    const Centre *centres;
    const sg_primitive *objects;
    const float4 *sky_light;     // 128 texels (src/shaders/ground.wgsl:10)
    const float *terrain;
    const sg_ray_star *catalogue;
    const uint32_t *sky_cells;   // 256*128+1 offsets, then indices
    const uint32_t *milky_way;
    const float4 *bvh;           // prop BVH nodes (2 float4 each) or null
    const uint32_t *bvh_prims;   // leaf slots -> primitive index
    TerrainLayout t;
    uint32_t catalogue_count;
    // Tabulated ring optical depth (SG_RING_PROFILE_BINS radial bins from
    // the inner to the outer edge) replacing the banded profile of one body.
    const float *ring_profile;
    uint32_t ring_profile_body;  // SG_MISS when no table is set
    // Tabulated relative air density (SG_DENSITY_BINS altitude slices up to
    // the top of the air) replacing exp(-h / H); null when not set.
    const float *density_profile;
    // 1 when a lit body has a solid-body style, 2-6 (kernels built with
    // SG_STYLES are required).
    uint32_t solid_styles;
    // Air march: view cells per camera ray (cap of the quarter-scale-height
    // rule) and density samples per light path.
    uint32_t air_view_steps;
    uint32_t air_sun_steps;
    // Mie scale height (AU). 0 keeps the reference's single scale height
    // (atmo_params.y) for both species; a positive value gives the Mie
    // density its own exp(-h / H_mie) while no density table is set.
    float mie_height;
};

struct Ray {
    d3 offset;        // relative to bodies[anchor] (or the camera when anchor == MISS)
    f3 direction;
    uint32_t anchor;
};
struct Hit {
    float distance;
    uint32_t index;
    f3 normal;
};
struct LightSample {
    f3 direction;
    float pdf;
};
struct RingHit {
    float distance;
    uint32_t index;
    f3 point;
};

SG_FN Hit miss_hit() { return Hit{1.0e30f, SG_MISS, f3{0.f, 0.f, 0.f}}; }

// --- RNG (src/shaders/pathtrace.wgsl:87-98) ------------------------------
SG_FN uint32_t hash_u32(uint32_t x) {
    x = (x ^ (x >> 16)) * 0x7feb352du;
    x = (x ^ (x >> 15)) * 0x846ca68bu;
    return x ^ (x >> 16);
}
SG_FN float rnd(uint32_t &state) {
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (float)(((word >> 22) ^ word) >> 8) * (1.0f / 16777216.0f);
}

// --- bodies and rebasing -----------------------------------------------------
SG_FN uint32_t body_count(const Scene &sc) { return sc.s.counts[0]; }
// Body records do not change while a kernel runs. Read through the constant
// address space, a uniform read is a scalar load in every kernel; through a
// plain global pointer it would be a vector load wherever a store of the
// same kernel may precede it (any read in a persistent loop).
SG_FN const Body &body(const Scene &sc, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstBody = __attribute__((address_space(4))) const Body;
    return *(const Body *)&((const ConstBody *)sc.bodies)[i];
#else
    return sc.bodies[i];
#endif
}
SG_FN d3 center_d(const Body &b) {
    return d3{(double)b.center.x + (double)b.low.x, (double)b.center.y + (double)b.low.y,
              (double)b.center.z + (double)b.low.z};
}
// The FP64 centres of `count` bodies (host side, at upload).
static inline void fill_centres(const Body *bodies, uint32_t count, Centre *out) {
    for (uint32_t i = 0; i < count; ++i) {
        const d3 c = center_d(bodies[i]);
        out[i] = Centre{c.x, c.y, c.z, 0.0};
    }
}
// Body i's FP64 centre, a uniform (scalar) load like body().
SG_FN d3 body_centre(const Scene &sc, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstCentre = __attribute__((address_space(4))) const Centre;
    const ConstCentre &c = ((const ConstCentre *)sc.centres)[i];
    return d3{c.x, c.y, c.z};
#else
    const Centre &c = sc.centres[i];
    return d3{c.x, c.y, c.z};
#endif
}
// src/shaders/pathtrace.wgsl:100-113, in FP64.
SG_FN d3 relative_center(const Scene &sc, const Ray &ray, uint32_t index) {
    if (ray.anchor == index) return -ray.offset;
    d3 c = body_centre(sc, index);
    if (ray.anchor == SG_MISS) return c - ray.offset;
    return (c - body_centre(sc, ray.anchor)) - ray.offset;
}

// src/shaders/pathtrace.wgsl:134-171: cross-product discriminant, one
// cancellation-free root and the product of roots for the other, normal kept
// body-local.
template <class T>
SG_FN Hit sphere_solve(V3<T> oc, V3<T> d, T radius, uint32_t index) {
    T a = dot(d, d);
    T b = dot(oc, d);
    V3<T> pc = cross(oc, d);
    T h = a * radius * radius - dot(pc, pc);
    if (h < T(0)) return miss_hit();
    T root = sqrtT(h);
    T q = b + (b >= T(0) ? root : -root);
    if (q == T(0)) return miss_hit();
    T c = dot(oc, oc) - radius * radius;
    T nearv = b < T(0) ? q / a : c / q;
    T farv = b < T(0) ? c / q : q / a;
    T distance = nearv;
    T delta = -root / a;
    if (distance <= T(0)) {
        distance = farv;
        delta = root / a;
    }
    if (distance <= T(0)) return miss_hit();
    V3<T> perpendicular = cross(d, pc) / a;
    V3<T> n = -perpendicular + d * delta;
    n = n * (T(1) / sqrtT(dot(n, n)));
    return Hit{(float)distance, index, f3{(float)n.x, (float)n.y, (float)n.z}};
}

SG_FN Hit sphere_hit_radius(const Scene &sc, const Ray &ray, uint32_t index, float radius) {
    d3 oc = relative_center(sc, ray, index);
    double r = radius;
    if (dot(oc, oc) < 16.0 * r * r) return sphere_solve<double>(oc, tod(ray.direction), r, index);
    return sphere_solve<float>(tof(oc), ray.direction, radius, index);
}
SG_FN Hit sphere_hit(const Scene &sc, const Ray &ray, uint32_t index) {
    SG_COUNT(SG_EV_SPHERE);
    return sphere_hit_radius(sc, ray, index, body(sc, index).center.w);
}

// src/shaders/pathtrace.wgsl:187-194
SG_FN f3 basis_direction(f3 axis, float x, float y, float z) {
    f3 helper = fabsf(axis.z) > 0.9f ? f3{0.f, 1.f, 0.f} : f3{0.f, 0.f, 1.f};
    f3 tangent = normalize(cross(helper, axis));
    return normalize(tangent * x + cross(axis, tangent) * y + axis * z);
}

// src/shaders/pathtrace.wgsl:197-202
SG_FN float cone_width(const Scene &sc, const Ray &ray, uint32_t index) {
    f3 oc = tof(relative_center(sc, ray, index));
    float r = body(sc, index).center.w;
    float sin2 = clampf(r * r / dot(oc, oc), 1.0e-20f, 1.0f);
    return sin2 / (1.0f + sqrtf(1.0f - sin2));
}
// src/shaders/pathtrace.wgsl:205-211
SG_FN uint32_t light_count(const Scene &sc) {
    return sc.s.lights[3] <= 3u ? sc.s.lights[3] : sc.s.counts[0];
}
SG_FN uint32_t light_index(const Scene &sc, uint32_t slot) {
    return sc.s.lights[3] <= 3u ? sc.s.lights[slot] : slot;
}
// src/shaders/pathtrace.wgsl:214-234
SG_FN LightSample sample_light_stratum(const Scene &sc, const Ray &ray, uint32_t index,
                                       uint32_t stratum, uint32_t count, uint32_t &rng) {
    f3 axis = normalize(tof(relative_center(sc, ray, index)));
    float width = cone_width(sc, ray, index);
    float radial = rnd(rng);
    float azimuth = rnd(rng);
    if (count == 4u) {
        radial = ((float)(stratum & 1u) + radial) * 0.5f;
        azimuth = ((float)(stratum >> 1) + azimuth) * 0.5f;
    }
    float one_minus_cos = radial * width;
    float sin_theta = sqrtf(one_minus_cos * (2.0f - one_minus_cos));
    float phi = TAU * azimuth;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return LightSample{basis_direction(axis, sin_theta * cp, sin_theta * sp, 1.0f - one_minus_cos),
                       1.0f / (TAU * width)};
}
SG_FN LightSample sample_light(const Scene &sc, const Ray &ray, uint32_t index, uint32_t &rng) {
    return sample_light_stratum(sc, ray, index, 0u, 1u, rng);
}
// src/shaders/pathtrace.wgsl:239-263. `accumulated` is counts.y of the
// dispatch the sample belongs to.
SG_FN uint32_t eclipse_light_samples(const Scene &sc, const Ray &ray, uint32_t index,
                                     uint32_t accumulated) {
    if (accumulated != 0u) return 1u;
    f3 source = tof(relative_center(sc, ray, index));
    float distance = length(source);
    f3 axis = source / distance;
    float source_chord = sqrtf(2.0f * cone_width(sc, ray, index));
    for (uint32_t b = 0; b < body_count(sc); ++b) {
        if (b == index) continue;
        f3 offset = tof(relative_center(sc, ray, b));
        float d2 = dot(offset, offset);
        float radius = body(sc, b).center.w;
        if (d2 <= radius * radius) continue;
        float d = sqrtf(d2);
        if (d - radius >= distance + body(sc, index).center.w) continue;
        float sin2 = clampf(radius * radius / d2, 0.0f, 1.0f);
        float limb_chord = sqrtf(2.0f * sin2 / (1.0f + sqrtf(1.0f - sin2)));
        float separation = length(axis - offset / d);
        if (separation < source_chord + limb_chord && separation + source_chord > limb_chord) return 4u;
    }
    return 1u;
}
// src/shaders/pathtrace.wgsl:264-267
SG_FN float power_weight(float a, float b) {
    float ratio = b / fmaxf(a, 1.0e-30f);
    return 1.0f / (1.0f + ratio * ratio);
}

// --- background (src/shaders/pathtrace.wgsl:115-128, 269-334) -------------
SG_FN f3 world_direction(const Scene &sc, f3 d) {
    const sg_globals &g = sc.s.globals;
    f3 r = {g.cam_right[0], g.cam_right[1], g.cam_right[2]};
    f3 u = {g.cam_up[0], g.cam_up[1], g.cam_up[2]};
    f3 f = {g.cam_forward[0], g.cam_forward[1], g.cam_forward[2]};
    return normalize(r * d.x + u * d.y - f * d.z);
}
SG_FN bool inside_star_dot(const Scene &sc, f3 view_direction, f3 star_direction) {
    const sg_globals &g = sc.s.globals;
    f3 r = {g.cam_right[0], g.cam_right[1], g.cam_right[2]};
    f3 u = {g.cam_up[0], g.cam_up[1], g.cam_up[2]};
    f3 f = {g.cam_forward[0], g.cam_forward[1], g.cam_forward[2]};
    f3 sv = {dot(star_direction, r), dot(star_direction, u), -dot(star_direction, f)};
    if (sv.z >= 0.0f) return false;
    float dx = view_direction.x / -view_direction.z - sv.x / -sv.z;
    float dy = view_direction.y / -view_direction.z - sv.y / -sv.z;
    float pixel_scale = g.viewport[1] / (2.0f * g.cam_forward[3]);
    dx *= pixel_scale;
    dy *= pixel_scale;
    return dx * dx + dy * dy <= STAR_DOT_PIXELS * STAR_DOT_PIXELS;
}
SG_FN f3 environment(const Scene &sc, f3 view_direction, bool primary) {
    f3 light = f3s(0.0f);
    if (!primary || sc.catalogue_count == 0u) return light;
    f3 direction = world_direction(sc, view_direction);
    float phi = atan2f(direction.y, direction.x) + PI;
    float theta = acosf(clampf(direction.z, -1.0f, 1.0f));
    uint32_t x = umin((uint32_t)(phi * (256.0f / TAU)), 255u);
    uint32_t y = umin((uint32_t)(theta * (128.0f / PI)), 127u);
    uint32_t cell = y * 256u + x;
    for (uint32_t j = sc.sky_cells[cell]; j < sc.sky_cells[cell + 1u]; ++j) {
        const sg_ray_star &star = sc.catalogue[sc.sky_cells[32769u + j]];
        f3 sd = {star.direction[0], star.direction[1], star.direction[2]};
        f3 perpendicular = cross(direction, sd);
        if (dot(direction, sd) > 0.0f && dot(perpendicular, perpendicular) <= star.direction[3]) {
            if (inside_star_dot(sc, view_direction, sd))
                light += f3{star.radiance[0], star.radiance[1], star.radiance[2]};
        }
    }
    return light;
}
SG_FN f3 milky_way_texel(const Scene &sc, int px, int py) {
    int width = (int)sc.s.background[1];
    int height = (int)sc.s.background[2];
    int x = (px + width) % width;
    int y = imin(imax(py, 0), height - 1);
    uint32_t packed = sc.milky_way[(uint32_t)(y * width + x)];
    return f3{(float)(packed & 1023u), (float)((packed >> 10) & 1023u), (float)((packed >> 20) & 1023u)} /
           1023.0f;
}
SG_FN f3 milky_way(const Scene &sc, f3 view_direction) {
    if (sc.s.background[0] == 0.0f || sc.milky_way == nullptr) return f3s(0.0f);
    f3 direction = world_direction(sc, view_direction);
    f3 pole = normalize(f3{0.35f, 0.82f, 0.45f});
    float a = -1.0f / (1.0f + pole.z);
    f3 right = {pole.x * pole.y * a, 1.0f + pole.y * pole.y * a, -pole.y};
    f3 forward = cross(pole, right);
    float longitude = atan2f(dot(direction, forward), dot(direction, right));
    float u = 0.5f - longitude / TAU;
    float v = acosf(clampf(dot(direction, pole), -1.0f, 1.0f)) / PI;
    float px = u * sc.s.background[1] - 0.5f;
    float py = v * sc.s.background[2] - 0.5f;
    int lx = (int)floorf(px), ly = (int)floorf(py);
    float wx = fractf(px), wy = fractf(py);
    f3 lum = mix3(mix3(milky_way_texel(sc, lx, ly), milky_way_texel(sc, lx + 1, ly), wx),
                  mix3(milky_way_texel(sc, lx, ly + 1), milky_way_texel(sc, lx + 1, ly + 1), wx), wy);
    return lum * sc.s.background[0];
}

// --- procedural albedo (src/shaders/pathtrace.wgsl:338-470, 897-925) ------
SG_FN float hash31(f3 p) {
    SG_EXACT
    // Scalar operations, so the pragma covers every one of them; the dot
    // product is two explicit fused multiply-adds in this order, as in the
    // oracle, so both evaluate the hash bit for bit.
    float qx = fractf(p.x * 0.1031f), qy = fractf(p.y * 0.1030f), qz = fractf(p.z * 0.0973f);
    const float d = fmaf(qx, qy + 33.33f, fmaf(qy, qx + 33.33f, qz * (qz + 33.33f)));
    qx += d;
    qy += d;
    qz += d;
    return fractf((qx + qy) * qz);
}
SG_FN float vnoise(f3 p) {
    f3 i = {floorf(p.x), floorf(p.y), floorf(p.z)};
    f3 f = p - i;
    f3 u = f * f * (f3s(3.0f) - f * 2.0f);
    float a = mixf(hash31(i), hash31(i + f3{1, 0, 0}), u.x);
    float b = mixf(hash31(i + f3{0, 1, 0}), hash31(i + f3{1, 1, 0}), u.x);
    float c = mixf(hash31(i + f3{0, 0, 1}), hash31(i + f3{1, 0, 1}), u.x);
    float d = mixf(hash31(i + f3{0, 1, 1}), hash31(i + f3{1, 1, 1}), u.x);
    return mixf(mixf(a, b, u.y), mixf(c, d, u.y), u.z);
}
SG_FN float fbm(f3 p, uint32_t octaves) {
    // Column-major mat3x3 of the WGSL: columns (0,1.6,1.2), (-1.6,.72,-.96), (-1.2,-.96,1.28).
    float amplitude = 0.5f, sum = 0.0f, norm = 0.0f;
    for (uint32_t o = 0; o < octaves; ++o) {
        sum += amplitude * vnoise(p);
        norm += amplitude;
        f3 r = {0.00f * p.x - 1.60f * p.y - 1.20f * p.z,
                1.60f * p.x + 0.72f * p.y - 0.96f * p.z,
                1.20f * p.x - 0.96f * p.y + 1.28f * p.z};
        p = r + f3{17.3f, 5.1f, 11.7f};
        amplitude *= 0.5f;
    }
    return sum / norm;
}
SG_FN f3 body_local(const Body &b, f3 n) {
    f3 x = xyz(b.axis_x), y = xyz(b.axis_y);
    return f3{dot(n, x), dot(n, y), dot(n, cross(x, y))};
}
SG_FN f3 seed_offset(float seed) {
    return f3{hash31(f3{seed, 1.0f, 7.0f}), hash31(f3{3.0f, seed, 5.0f}), hash31(f3{9.0f, 2.0f, seed})} *
           512.0f;
}
// The solid-body albedos (styles 2-6) are called out of line on the GPU,
// which keeps their crater lists and loops out of the callers' register
// allocation.
#if defined(__HIP_DEVICE_COMPILE__)
#define SG_OUTLINE static __device__ __attribute__((noinline))
#else
#define SG_OUTLINE static inline
#endif
// src/shaders/pathtrace.wgsl:394-470
// The gas giant's colour is a chain of component-wise steps in its zone and
// belt colours (material.rgb, surface_color.rgb) whose weights depend only
// on the point: gas_giant_pattern finds the weights, giant_channel runs the
// chain for one channel, in plain floats for the albedo and in value and
// derivative pairs for d albedo / d zone.
struct GiantPattern {
    float bands, tone, shade, fine, polar;
    uint32_t storms;
    float masks[3], rims[3], cores[3];
};
SG_FN GiantPattern gas_giant_pattern(const Body &b, f3 local) {
    f3 off = seed_offset(b.axis_y.w);
    float turbulence = b.surface_color.w;
    float bands = b.surface_params.x;
    uint32_t storms = (uint32_t)(b.surface_params.y + 0.5f);
    f3 q = local;
    GiantPattern g;
    g.storms = storms < 3u ? storms : 3u;
    for (uint32_t k = 0; k < 3u; ++k) g.masks[k] = g.rims[k] = 0.0f;
    for (uint32_t k = 0; k < storms && k < 3u; ++k) {
        float fk = (float)k;
        float lat = (hash31(off + f3{7.1f * fk, 1.3f, 2.9f}) - 0.5f) * 1.1f;
        float lon = hash31(off + f3{3.7f, 5.3f * fk, 8.1f}) * TAU;
        float size = (k == 0u ? 0.11f : 0.05f) * (0.8f + 0.4f * hash31(off + f3{fk, 9.0f, 4.0f}));
        f3 centre = {cosf(lat) * cosf(lon), cosf(lat) * sinf(lon), sinf(lat)};
        if (dot(q, centre) <= 0.0f) continue;
        f3 east = {-sinf(lon), cosf(lon), 0.0f};
        f3 north = cross(centre, east);
        float dx = dot(q - centre, east) / (2.0f * size);
        float dy = dot(q - centre, north) / size;
        float e = sqrtf(dx * dx + dy * dy);
        float influence = 1.0f - smoothstepf(0.0f, 2.4f, e);
        float angle = (lat < 0.0f ? 3.2f : -3.2f) * influence * influence;
        float c = cosf(angle), s = sinf(angle);
        float rx = c * dx - s * dy;
        float ry = s * dx + c * dy;
        q = normalize(q + east * ((rx - dx) * 2.0f * size) + north * ((ry - dy) * size));
        float ragged = e + 0.25f * (fbm(q * 18.0f + off, 3) - 0.5f);
        g.masks[k] = 1.0f - smoothstepf(0.7f, 1.0f, ragged);
        g.rims[k] = smoothstepf(0.75f, 1.0f, ragged) * (1.0f - smoothstepf(1.0f, 1.35f, ragged));
    }
    f3 p = f3{q.x * 2.2f, q.y * 2.2f, q.z * 9.0f} + off;
    float warp1 = fbm(p, 5) - 0.5f;
    float warp2 = fbm(p * 2.1f + f3{5.2f, 1.3f, 7.7f} + f3s(4.0f * warp1), 5) - 0.5f;
    float lat = q.z + 0.07f * sinf(3.0f * PI * q.z + off.z) + 0.04f * sinf(5.0f * PI * q.z + off.y);
    float shear = 0.35f + 0.65f * fabsf(cosf(PI * bands * lat + off.x));
    float lon = atan2f(q.y, q.x);
    float ripple = sinf(lon * (9.0f + floorf(bands)) + 6.0f * warp1 + off.y) * 0.5f +
                   sinf(lon * 23.0f + 9.0f * warp2 + off.z) * 0.25f;
    float mu = lat + turbulence * shear * (0.16f * warp1 + 0.10f * warp2 + 0.025f * ripple);
    float wave = 0.5f + 0.5f * sinf(PI * bands * mu + off.x);
    float irregular = fbm(f3{mu * bands * 1.7f, off.y, off.z}, 3);
    float profile = 0.5f * wave + 0.5f * irregular;
    float tone = fbm(f3{mu * bands * 0.9f, off.z, off.x}, 2);
    float fine = fbm(f3{q.x * 7.0f, q.y * 7.0f, q.z * 70.0f} + off + f3s(3.0f * warp2), 4);
    float eddies = fbm(p * 5.0f + f3{warp1, warp2, warp1} * 6.0f, 4);
    g.bands = smoothstepf(0.3f, 0.75f, profile);
    g.tone = smoothstepf(0.35f, 0.75f, tone) * 0.8f;
    g.shade = (0.78f + 0.44f * fine) * (0.9f + 0.2f * eddies);
    g.fine = fine;
    g.polar = smoothstepf(0.55f, 0.92f, fabsf(local.z)) * 0.8f;
    for (uint32_t k = 0; k < 3u; ++k)
        g.cores[k] = 1.0f - 0.25f * (1.0f - smoothstepf(0.0f, 0.6f, g.masks[k] - g.rims[k]));
    return g;
}
// A value with its derivative along one parameter.
struct Dual {
    float v, d;
};
SG_FN Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
SG_FN Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
SG_FN Dual operator*(Dual a, float s) { return {a.v * s, a.d * s}; }
SG_FN Dual operator+(Dual a, float s) { return {a.v + s, a.d}; }
SG_FN Dual mix_t(Dual a, Dual b, float t) { return a * (1.0f - t) + b * t; }
SG_FN float mix_t(float a, float b, float t) { return a * (1.0f - t) + b * t; }
template <class T>
SG_FN T giant_channel(const GiantPattern &g, T zone, T belt) {
    T color = mix_t(belt, zone, g.bands);
    color = mix_t(color, color * (color * 0.5f + 0.55f), g.tone);
    color = color * g.shade;
    color = mix_t(color, mix_t(zone, belt, 0.55f) * (0.8f + 0.2f * g.fine), g.polar);
    for (uint32_t k = 0; k < g.storms; ++k) {
        T tint = k == 0u ? belt * belt * 1.5f : zone * 1.15f;
        color = mix_t(color, tint * (0.7f + 0.6f * g.fine) * g.cores[k], g.masks[k] * 0.92f);
        color = mix_t(color, zone * 1.1f, g.rims[k] * 0.45f);
    }
    return color;
}
SG_FN f3 gas_giant_albedo(const Body &b, f3 local) {
    const GiantPattern g = gas_giant_pattern(b, local);
    return f3{giant_channel(g, b.material.x, b.surface_color.x), giant_channel(g, b.material.y, b.surface_color.y),
              giant_channel(g, b.material.z, b.surface_color.z)};
}
SG_FN uint32_t surface_style(const Body &b) { return (uint32_t)(b.axis_x.w + 0.5f); }
// --- Solid-body surfaces (src/shaders/pathtrace.wgsl:472-689) ------------
// wgsl:477-479
SG_FN float surface_gaussian(float d) { return expf(-(d * d)); }
// wgsl:481-484
SG_FN f3 hash33(f3 p) {
    return f3{hash31(p), hash31(p + f3{19.1f, 7.3f, 3.7f}), hash31(p + f3{5.9f, 31.7f, 11.3f})};
}
// Dark lava plains in [0, 1] (wgsl:492-495).
SG_FN float maria(f3 q, f3 off) {
    const float warp = fbm(q * 2.0f + off, 4u) - 0.5f;
    return smoothstepf(0.5f, 0.56f, fbm(q * 1.2f + off + f3s(1.5f * warp), 5u));
}
// The crater owned by grid cell `cell` (wgsl:502-511): a jittered point of
// the cell, projected onto the unit sphere, kept if the projection stays in
// the cell; three jitters are tried. w = 0 when the cell has none.
SG_FN float4 cell_centre(f3 cell, f3 key, float scale, f3 off) {
    for (uint32_t t = 0; t < 3u; ++t) {
        const f3 c = normalize(cell + hash33(key + f3s((float)t * 17.0f)) - off);
        const f3 g = c * scale + off;
        if (floorf(g.x) == cell.x && floorf(g.y) == cell.y && floorf(g.z) == cell.z)
            return make_float4(c.x, c.y, c.z, 1.0f);
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
SG_FN f3 crater_cell(f3 base, uint32_t i) {
    return base + f3{(float)(i & 1u), (float)((i >> 1) & 1u), (float)((i >> 2) & 1u)};
}
// One octave of craters (wgsl:527-633). state.x is the floor darkness (a
// multiplier), state.y the relief brightness (additive). The craters that
// reach q (two slots per cell, the 2x2x2 nearest cells) are painted oldest
// first; at most the six youngest are kept.
SG_FN float2 crater_layer(f3 q, float scale, float density, f3 off, f3 body_off, float flood, float contrast,
                          float rays, float2 state) {
    const f3 p = q * scale + off;
    const f3 base = {floorf(p.x - 0.5f), floorf(p.y - 0.5f), floorf(p.z - 0.5f)};
    uint32_t slots[6];
    float ages[6];
    f3 centres[6];
    uint32_t count = 0;
    for (uint32_t i = 0; i < 16u; ++i) {
        const f3 grid = crater_cell(base, i);
        const f3 gap = {fmaxf(fmaxf(grid.x - p.x, p.x - grid.x - 1.0f), 0.0f),
                        fmaxf(fmaxf(grid.y - p.y, p.y - grid.y - 1.0f), 0.0f),
                        fmaxf(fmaxf(grid.z - p.z, p.z - grid.z - 1.0f), 0.0f)};
        if (dot(gap, gap) > 0.2343f) continue;
        const f3 key = grid + f3s((float)(i >> 3) * 1000.0f);
        if (hash31(key + f3{71.3f, 3.1f, 9.7f}) > density) continue;
        const float4 found = cell_centre(grid, key, scale, off);
        if (found.w == 0.0f) continue;
        const f3 centre = {found.x, found.y, found.z};
        const float size = hash31(key + f3{3.1f, 8.3f, 1.7f});
        const float radius = (0.06f + 0.16f * size * size) / scale;
        if (length(q - centre) > 2.2f * radius) continue;
        const float age = hash31(key + f3{29.0f, 61.0f, 7.0f});
        if (flood > 0.0f && age < 0.7f && maria(centre, body_off) > 0.5f) continue;
        if (count == 6u) {
            if (age <= ages[0]) continue;
            for (uint32_t m = 0; m < 5u; ++m) {
                ages[m] = ages[m + 1];
                slots[m] = slots[m + 1];
                centres[m] = centres[m + 1];
            }
            count = 5u;
        }
        uint32_t j = count;
        while (j > 0u && ages[j - 1] > age) {
            ages[j] = ages[j - 1];
            slots[j] = slots[j - 1];
            centres[j] = centres[j - 1];
            --j;
        }
        ages[j] = age;
        slots[j] = i;
        centres[j] = centre;
        ++count;
    }
    float floor_shade = 1.0f;
    float relief = state.y;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t i = slots[k];
        const f3 cell = crater_cell(base, i) + f3s((float)(i >> 3) * 1000.0f);
        const f3 centre = centres[k];
        const float size = hash31(cell + f3{3.1f, 8.3f, 1.7f});
        const float radius = (0.06f + 0.16f * size * size) / scale;
        const f3 offset = q - centre;
        const float d = length(offset) / radius;
        f3 east = cross(f3{0.0f, 0.0f, 1.0f}, centre);
        east = east / fmaxf(length(east), 1.0e-4f);
        const f3 north = cross(centre, east);
        const float angle = atan2f(dot(offset, north), dot(offset, east));
        const f3 phase = hash33(cell + f3{41.0f, 5.0f, 23.0f}) * TAU;
        const float wobble = 0.025f * (sinf(3.0f * angle + phase.x) + sinf(5.0f * angle + phase.y) +
                                       0.6f * sinf(8.0f * angle + phase.z));
        const float r = d + wobble;
        floor_shade = fminf(floor_shade, 1.0f - 0.18f * contrast * (1.0f - smoothstepf(0.55f, 0.9f, r)));
        const float rim = 0.22f * contrast * surface_gaussian((r - 1.0f) / 0.13f);
        const float footprint = 1.0f - smoothstepf(1.1f, 1.3f, r);
        relief = mixf(relief, rim, footprint);
        const float blanket = (1.0f - smoothstepf(1.0f, 1.9f, r)) * (1.0f - footprint);
        relief = mixf(relief, 0.0f, blanket * 0.3f) + 0.1f * contrast * blanket;
        const float fresh = hash31(cell + f3{13.0f, 2.0f, 5.0f});
        if (rays > 0.0f && fresh > 0.82f) {
            const float n = vnoise(f3{cosf(angle) * 9.0f, sinf(angle) * 9.0f, fresh * 40.0f});
            const float streaks = n * n * n;
            relief += contrast * (rays * 0.9f * streaks * (1.0f - smoothstepf(1.0f, 2.2f, d)) * (d >= 0.9f ? 1.0f : 0.0f) +
                                  rays * 0.25f * (1.0f - smoothstepf(0.0f, 1.0f, d)));
        }
    }
    return make_float2(state.x * floor_shade, relief);
}
// Five octaves, large (oldest) to small (wgsl:636-647).
SG_FN float craters(f3 q, float density, f3 off, float flood, float contrast, float rays) {
    if (density <= 0.0f) return 1.0f;
    float2 st = crater_layer(q, 1.5f, density, off, off, flood, contrast, rays, make_float2(1.0f, 0.0f));
    st = crater_layer(q, 3.5f, density, off + f3s(101.0f), off, flood, contrast, rays * 0.8f, st);
    st = crater_layer(q, 8.0f, density, off + f3s(203.0f), off, flood, contrast * 0.85f, rays * 0.5f, st);
    st = crater_layer(q, 18.0f, density, off + f3s(307.0f), off, 0.0f, contrast * 0.7f, 0.0f, st);
    st = crater_layer(q, 40.0f, density, off + f3s(401.0f), off, 0.0f, contrast * 0.55f, 0.0f, st);
    return st.x * (1.0f + st.y);
}
// Polar ice down to |sin(latitude)| = 1 - extent, ragged (wgsl:650-656).
SG_FN float polar_caps(f3 q, float extent, f3 off) {
    if (extent <= 0.0f) return 0.0f;
    const float edge = 1.0f - extent + 0.12f * (fbm(q * 5.0f + off, 5u) - 0.5f) + 0.05f * (fbm(q * 18.0f + off, 3u) - 0.5f);
    return smoothstepf(edge - 0.015f, edge + 0.015f, fabsf(q.z));
}
// Moon-like: highlands, flooded maria, craters, regolith grain (wgsl:660-670).
SG_OUTLINE f3 cratered_albedo(const Body &b, f3 q) {
    const f3 off = seed_offset(b.axis_y.w);
    const float mare = maria(q, off);
    const float grain = fbm(q * 45.0f + off, 3u);
    f3 color = mix3(xyz(b.material), xyz(b.surface_color), mare);
    color = color * (0.9f + 0.2f * grain);
    return color * craters(q, b.surface_color.w, off, 1.0f, 1.0f, 1.0f);
}
// Mars-like: ochre plains, dark regions, soft craters, streaks, polar caps
// (wgsl:673-689).
SG_OUTLINE f3 rocky_albedo(const Body &b, f3 q) {
    const f3 off = seed_offset(b.axis_y.w);
    const float wx = fbm(q * 1.8f + off, 4u) - 0.5f, wy = fbm(q * 1.8f + off + f3s(7.0f), 4u) - 0.5f;
    const float dark = smoothstepf(0.5f, 0.64f, fbm(q * 2.2f + off + f3{2.0f * wx, 2.0f * wy, wx}, 6u));
    const float tone = fbm(q * 5.0f + off + f3s(3.0f), 4u);
    f3 color = mix3(xyz(b.material), xyz(b.surface_color), dark * 0.85f);
    color = color * (0.85f + 0.3f * tone);
    const float streak = fbm(f3{q.x * 4.0f, q.y * 4.0f, q.z * 18.0f} + off + f3s(9.0f), 3u);
    color = color * (0.95f + 0.1f * streak);
    color = color * craters(q, b.surface_color.w, off, 0.0f, 0.45f, 0.0f);
    const float ice = polar_caps(q, b.surface_params.x, off);
    return mix3(color, f3{0.92f, 0.92f, 0.94f}, ice);
}
// Thin bright lines where the noise crosses 0.5 (wgsl:487-489).
SG_FN float ridge(float n, float width) { return 1.0f - smoothstepf(0.0f, width, fabsf(n - 0.5f)); }
// Nearest-cell distance, gap to the second nearest, and the nearest plate's
// tone over all 27 neighbour cells (wgsl:691-709).
SG_FN f3 ice_plates(f3 p) {
    const f3 base = {floorf(p.x), floorf(p.y), floorf(p.z)};
    float nearest = 10.0f, second = 10.0f, tone = 0.0f;
    for (int i = 0; i < 27; ++i) {
        const f3 cell = base + f3{(float)(i % 3 - 1), (float)((i / 3) % 3 - 1), (float)(i / 9 - 1)};
        const f3 delta = cell + f3s(0.2f) + hash33(cell) * 0.6f - p;
        const float d = dot(delta, delta);
        if (d < nearest) {
            second = nearest;
            nearest = d;
            tone = hash31(cell + f3s(57.0f));
        } else {
            second = fminf(second, d);
        }
    }
    return f3{sqrtf(nearest), sqrtf(second) - sqrtf(nearest), tone};
}
// Europa-like: disrupted ice, plates, lineae in fracture families, sparse
// craters (wgsl:714-767).
SG_OUTLINE f3 icy_albedo(const Body &b, f3 q) {
    const f3 off = seed_offset(b.axis_y.w);
    const f3 ice = xyz(b.material);
    const f3 rust = xyz(b.surface_color) * f3{0.72f, 0.62f, 0.56f};
    const f3 warp = f3{fbm(q * 3.1f + off + f3s(11.0f), 3u), fbm(q * 3.1f + off + f3s(23.0f), 3u),
                       fbm(q * 3.1f + off + f3s(37.0f), 3u)} -
                    f3s(0.5f);
    const float terrain = fbm(q * 4.2f + off + warp, 5u);
    const float rubble = fbm(q * 85.0f + off + warp * 7.0f, 4u);
    const float chaos = smoothstepf(0.42f, 0.65f, terrain + 0.14f * (rubble - 0.5f));
    const f3 plates = ice_plates(q * 72.0f + off + warp * 4.0f);
    const float seams = 1.0f - smoothstepf(0.015f, 0.07f, plates.y);
    const float dust = chaos * (0.2f + 0.46f * rubble + 0.18f * plates.z);
    f3 color = mix3(ice, rust, dust);
    color = mix3(color, rust * 0.8f, seams * chaos * 0.35f);
    const float ridged_plain = ridge(fbm(q * 34.0f + off + f3s(83.0f), 3u), 0.025f);
    color = mix3(color, rust, ridged_plain * (0.1f + 0.14f * chaos));
    color = color * (0.87f + 0.22f * fbm(q * 180.0f + off, 3u));
    const f3 warped = q + warp * 0.1f;
    const float roughness = 0.0015f * (rubble - 0.5f);
    float shoulder = 0.0f, fissure = 0.0f, fresh_ridge = 0.0f;
    for (uint32_t i = 0; i < 64u; ++i) {
        const f3 h = hash33(off + f3{(float)i * 3.7f, 1.0f, 2.0f});
        const f3 family = hash33(off + f3{(float)(i / 8u) * 17.3f, 21.0f, 7.0f});
        const f3 n = normalize(family * 2.0f - f3s(1.0f) + (h * 2.0f - f3s(1.0f)) * 0.28f);
        const float shift = (hash31(off + f3{(float)i, 9.0f, 4.0f}) - 0.5f) * 1.4f;
        const bool old_band = i < 10u;
        const float width = (old_band ? 0.0015f + 0.0022f * h.z : 0.0005f + 0.0011f * h.z) *
                            (0.65f + 0.8f * vnoise(q * 17.0f + h * 29.0f));
        const float bend = 0.03f * (vnoise(q * 6.0f + family * 31.0f) - 0.5f) +
                           0.008f * (vnoise(q * 24.0f + family * 37.0f) - 0.5f);
        const float dist = fabsf(dot(warped, n) - shift + roughness + bend);
        const float along = smoothstepf(0.28f, 0.62f, vnoise(q * 3.5f + h * 50.0f));
        shoulder = fmaxf(shoulder, (1.0f - smoothstepf(width, width * 4.0f, dist)) * along);
        const float double_ridge = fabsf(dist - width * 0.85f);
        const float single = 1.0f - smoothstepf(width * 0.25f, width, dist);
        const float paired = 1.0f - smoothstepf(width * 0.15f, width * 0.55f, double_ridge);
        fissure = fmaxf(fissure, (old_band ? paired : single) * along);
        fresh_ridge = fmaxf(fresh_ridge,
                            (1.0f - smoothstepf(width * 0.15f, width * 0.35f, fabsf(dist - width * 1.8f))) * along);
    }
    color = mix3(color, rust, shoulder * 0.17f);
    color = mix3(color, rust * 0.7f, fissure * 0.8f);
    color = mix3(color, ice, fresh_ridge * 0.13f);
    return color * craters(q, b.surface_color.w * 0.2f, off, 0.0f, 0.3f, 0.4f);
}
// Io-like: sulphur plains, frost, paterae, lava lobes, plume deposits over
// the 27 cells around the point (wgsl:770-851).
SG_OUTLINE f3 volcanic_albedo(const Body &b, f3 q) {
    const f3 off = seed_offset(b.axis_y.w);
    const f3 sulphur = xyz(b.material);
    const f3 deposit = xyz(b.surface_color);
    const f3 warp = f3{fbm(q * 4.7f + off, 3u), fbm(q * 4.7f + off + f3s(19.0f), 3u),
                       fbm(q * 4.7f + off + f3s(41.0f), 3u)} -
                    f3s(0.5f);
    const float terrain = fbm(q * 10.0f + off + warp * 2.0f, 6u);
    const float weathered = smoothstepf(0.38f, 0.65f, fbm(q * 3.7f + off + f3s(11.0f), 5u));
    const float fine = fbm(q * 90.0f + off + warp * 3.0f, 4u);
    const f3 crust = mix3(sulphur * f3{0.43f, 0.51f, 0.6f}, deposit, 0.12f);
    f3 color = sulphur * (0.7f + 0.5f * terrain);
    color = mix3(color, crust, weathered * (0.45f + 0.5f * terrain));
    const float frost = smoothstepf(0.38f, 0.74f, fbm(q * 6.0f + off + f3s(29.0f) + warp, 5u) + 0.2f * (fine - 0.5f));
    const f3 frost_color = mix3(sulphur, f3{0.72f, 0.7f, 0.62f}, 0.8f);
    color = mix3(color, frost_color, frost * 0.7f);
    const float flow_edges = ridge(fbm(q * 35.0f + off + warp * 2.0f, 4u), 0.025f);
    color = mix3(color, crust, flow_edges * weathered * 0.4f);
    color = color * ((0.82f + 0.32f * fine) * (0.94f + 0.13f * fbm(q * 210.0f + off, 3u)));
    float dark = 0.0f, ash = 0.0f, red = 0.0f, white = 0.0f;
    const float density = b.surface_color.w;
    const float scale = 5.5f;
    const f3 p = q * scale + off + f3s(300.0f);
    const f3 base = {floorf(p.x), floorf(p.y), floorf(p.z)};
    for (int i = 0; i < 27; ++i) {
        const f3 cell = base + f3{(float)(i % 3 - 1), (float)((i / 3) % 3 - 1), (float)(i / 9 - 1)};
        if (hash31(cell + f3{4.0f, 41.0f, 9.0f}) > 0.6f * density) continue;
        const float4 found = cell_centre(cell, cell, scale, off + f3s(300.0f));
        if (found.w == 0.0f) continue;
        const f3 centre = {found.x, found.y, found.z};
        const float size = hash31(cell + f3{8.0f, 1.0f, 3.0f});
        const float radius = (0.025f + 0.11f * size * size * size) / scale;
        const f3 offset = q - centre;
        const float reach = length(offset) * scale;
        if (reach > 0.95f) continue;
        const float cutoff = 1.0f - smoothstepf(0.75f, 0.95f, reach);
        const f3 tangent = cross(centre, f3{0.37f, 0.83f, 0.41f});
        const f3 east = tangent / fmaxf(length(tangent), 1.0e-5f);
        const f3 north = cross(centre, east);
        const float angle = TAU * hash31(cell + f3s(91.0f));
        const float ux = dot(offset, east) / radius, uy = dot(offset, north) / radius;
        const float ca = cosf(angle), sa = sinf(angle);
        const float vx = ux * ca - uy * sa, vy = ux * sa + uy * ca;
        const float ragged = fbm(f3{vx * 2.7f, vy * 2.7f, 0.0f} + cell * 7.0f, 3u) - 0.5f;
        const float ex = vx * (1.0f / (0.65f + 0.9f * size));
        const float d = sqrtf(ex * ex + vy * vy) + 0.42f * ragged;
        const float pit = 1.0f - smoothstepf(0.68f, 0.86f, d);
        const float age = hash31(cell + f3s(31.0f));
        const float fx = vx - (0.2f + 2.3f * age), fy = vy - 0.35f * sinf(vx * 1.3f + angle);
        const float sx = fx * (0.75f - 0.48f * age), sy = fy * (0.85f + 0.35f * size);
        const float flow_shape = sqrtf(sx * sx + sy * sy) + 0.65f * ragged +
                                 0.35f * (fbm(f3{vx * 1.8f, vy * 1.8f, 0.0f} + cell * 3.0f, 4u) - 0.5f);
        const float flow = (1.0f - smoothstepf(0.72f, 0.9f, flow_shape)) * smoothstepf(0.35f, 0.85f, age) * 0.85f;
        dark = fmaxf(dark, fmaxf(pit, flow) * cutoff);
        const float plume = hash31(cell + f3{2.0f, 6.0f, 17.0f});
        const float hx = (vx - 0.5f) * 0.16f, hy = (vy - 0.3f) * 0.23f;
        const float halo = expf(-(hx * hx + hy * hy)) * cutoff;
        ash = fmaxf(ash, halo * (0.25f + 0.35f * age));
        white = fmaxf(white, halo * (plume >= 0.45f ? 1.0f : 0.0f) * (1.0f - (plume >= 0.83f ? 1.0f : 0.0f)));
        const float apron = surface_gaussian((sqrtf(vx * vx + vy * vy) + ragged - 3.5f) / 1.6f) *
                            (0.25f + 0.75f * smoothstepf(-1.5f, 2.0f, vy));
        red = fmaxf(red, apron * (plume >= 0.83f ? 1.0f : 0.0f) * cutoff);
    }
    color = mix3(color, crust * 0.7f, ash * 0.45f);
    color = mix3(color, frost_color, white * 0.4f);
    color = mix3(color, deposit * f3{0.85f, 0.7f, 0.65f}, red * 0.6f);
    return mix3(color, f3{0.035f, 0.032f, 0.026f}, dark * 0.96f);
}
// Earth-like: oceans and shallows, continents by moisture and altitude,
// polar ice, cloud belts (wgsl:854-895).
SG_OUTLINE f3 terran_albedo(const Body &b, f3 q) {
    const f3 off = seed_offset(b.axis_y.w);
    const f3 land_base = xyz(b.material);
    const f3 ocean_color = xyz(b.surface_color);
    const f3 warp = f3{fbm(q * 1.5f + off, 4u), fbm(q * 1.5f + off + f3s(13.0f), 4u),
                       fbm(q * 1.5f + off + f3s(29.0f), 4u)} -
                    f3s(0.5f);
    const float height = fbm(q * 1.1f + off + warp * 0.9f, 7u);
    const float sea = 0.5f + (b.surface_params.y - 0.5f) * 0.3f;
    const float coast = smoothstepf(sea, sea + 0.004f, height);
    const float shallow = 1.0f - smoothstepf(sea - 0.04f, sea, height);
    const f3 water = mix3(ocean_color * 1.8f + f3{0.0f, 0.04f, 0.05f}, ocean_color, shallow);
    const float moisture = fbm(q * 3.0f + off + f3s(41.0f), 5u) + 0.25f * (fabsf(q.z) - 0.35f);
    const f3 green = land_base * f3{0.45f, 0.62f, 0.38f};
    f3 land = mix3(land_base, green, smoothstepf(0.45f, 0.6f, moisture));
    const float altitude = smoothstepf(sea + 0.08f, sea + 0.2f, height);
    land = mix3(land, f3{0.42f, 0.38f, 0.34f}, altitude * 0.6f);
    land = mix3(land, f3{0.9f, 0.9f, 0.92f}, smoothstepf(sea + 0.19f, sea + 0.24f, height));
    land = land * (0.88f + 0.24f * fbm(q * 25.0f + off, 3u));
    f3 color = mix3(water, land, coast);
    const float ice = polar_caps(q, b.surface_params.x, off);
    color = mix3(color, f3{0.9f, 0.92f, 0.95f}, ice);
    const float cover = b.surface_params.z;
    if (cover > 0.0f) {
        const f3 w1 = f3{fbm(q * 1.6f + off + f3s(77.0f), 4u), fbm(q * 1.6f + off + f3s(91.0f), 4u), 0.5f} - f3s(0.5f);
        const f3 w2 = f3{fbm(q * 4.0f + off + w1 * 3.0f, 4u), fbm(q * 4.0f + off + f3s(17.0f) + w1 * 3.0f, 4u), 0.5f} -
                      f3s(0.5f);
        const f3 cq = f3{q.x * 2.5f, q.y * 2.5f, q.z * 7.0f} + off + w1 * 2.5f + w2 * 1.2f;
        const float cloud = fbm(cq, 7u);
        const float lat = fabsf(q.z);
        const float belts = 0.06f * (surface_gaussian(lat / 0.12f) + surface_gaussian((lat - 0.72f) / 0.15f)) -
                            0.04f * surface_gaussian((lat - 0.42f) / 0.1f);
        const float threshold = 0.62f - 0.22f * cover - belts;
        const float mask = smoothstepf(threshold - 0.04f, threshold + 0.14f, cloud);
        color = mix3(color, f3{0.93f, 0.94f, 0.96f}, mask * 0.95f);
    }
    return color;
}
// The unclamped albedo of procedural styles 1-6 at the local point q.
// Built without SG_STYLES, a kernel serves frames whose bodies use styles 0
// and 1 only (the host library selects the kernels). Codes above 6 are
// terran, as in the WGSL switch.
SG_FN f3 style_albedo(const Body &b, f3 q) {
#if SG_STYLES
    switch (surface_style(b)) {
    case 1u: return gas_giant_albedo(b, q);
    case 2u: return cratered_albedo(b, q);
    case 3u: return rocky_albedo(b, q);
    case 4u: return icy_albedo(b, q);
    case 5u: return volcanic_albedo(b, q);
    default: return terran_albedo(b, q);
    }
#else
    return gas_giant_albedo(b, q);
#endif
}
// src/shaders/pathtrace.wgsl:897-925 for styles 0 (mottle) and 1-6 (gas
// giant, cratered, rocky, icy, volcanic, terran).
// Style 0: the albedo is material.rgb times this scalar mottle, clamped.
SG_FN float surface_mottle(const Body &b, f3 local) {
    float mottle = 1.0f;
    if (b.low.w > 0.0f) {
        f3 p = local * 2.4f + seed_offset(b.axis_y.w);
        float amplitude = 0.5f, noise = 0.0f;
        for (uint32_t o = 0; o < 5u; ++o) {
            noise += amplitude * vnoise(p);
            p = p * 2.07f;
            amplitude *= 0.5f;
        }
        mottle = mixf(1.0f, 0.68f + 0.64f * noise, b.low.w);
    }
    return mottle;
}
SG_FN f3 surface_albedo(const Body &b, f3 normal) {
    f3 local = body_local(b, normal);
    if (surface_style(b) > 0u) return clamp3(style_albedo(b, local), 0.0f, 1.0f);
    return clamp3(xyz(b.material) * surface_mottle(b, local), 0.0f, 1.0f);
}
// surface_albedo and, in `derivative`, d albedo / d material.rgb for styles 0
// and 1 (zero where clamped, and for the solid-body styles 2-6, which the
// adjoint rejects). The gas giant's pattern is found once for both.
SG_FN f3 surface_albedo_with_derivative(const Body &b, f3 normal, f3 &derivative) {
    const f3 local = body_local(b, normal);
    const uint32_t style = surface_style(b);
    f3 a, d = f3s(0.0f);
    if (style == 1u) {
        const GiantPattern g = gas_giant_pattern(b, local);
        const Dual x = giant_channel(g, Dual{b.material.x, 1.0f}, Dual{b.surface_color.x, 0.0f});
        const Dual y = giant_channel(g, Dual{b.material.y, 1.0f}, Dual{b.surface_color.y, 0.0f});
        const Dual z = giant_channel(g, Dual{b.material.z, 1.0f}, Dual{b.surface_color.z, 0.0f});
        a = f3{x.v, y.v, z.v};
        d = f3{x.d, y.d, z.d};
    } else if (style == 0u) {
        const float m = surface_mottle(b, local);
        a = xyz(b.material) * m;
        d = f3s(m);
    } else {
        a = style_albedo(b, local);
    }
    derivative = f3{(a.x > 0.0f && a.x < 1.0f) ? d.x : 0.0f, (a.y > 0.0f && a.y < 1.0f) ? d.y : 0.0f,
                    (a.z > 0.0f && a.z < 1.0f) ? d.z : 0.0f};
    return clamp3(a, 0.0f, 1.0f);
}

}  // namespace sg
