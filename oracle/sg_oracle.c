/*
 * ORACLE — test infrastructure only (see sg_oracle.h).
 *
 * CPU restatement of stargaze's WGSL path tracer, line by line, in plain C.
 * Built twice:
 *   REAL=float  with -fsingle-precision-constant -ffp-contract=off: the
 *               reference's FP32 arithmetic as written (also the CPU baseline);
 *   REAL=double: an FP64 evaluation of the same algorithm.
 *
 * Each function cites the WGSL it restates (paths relative to the reference
 * repository). Transcendentals are libm's; the reference's are driver-defined
 * within WGSL bounds, so bit-exactness with any GPU is not a goal — the
 * restatement is pinned instead by the reference's own closed-form and
 * brute-force tests, restated in tests/test_oracle_golden.py:
 *   hash/random ........... PCG-RXS-M-XS and lowbias32 known answers
 *   sphere_hit ............ emitter radiance, inside-sphere roots, observer 2 m
 *                           above the host, tiny distant spheres
 *   sample_light/MIS ...... analytic irradiance, binary suns, MIS ±5%
 *   light_visibility ...... umbra / penumbra / annular, occluder behind emitter
 *   ring_hit/optics/bsdf .. t, radius and Beer–Lambert closed forms, slab
 *                           single scattering, energy <= 1
 *   atmosphere ............ zenith extinction x/(2 sinh(x/2)), eclipse = 0
 *   terrain_query ......... exhaustive f64 triangles
 *   props ................. box / ellipsoid closed forms
 *   cratered/rocky albedo . the material probe (finite, non-negative, never
 *                           black), polar-cap and crater-floor closed forms
 *                           (tests/test_surface_styles.py)
 * Functions without such a pin are listed as "parity unpinned" in the
 * oracle's documentation: gas_giant_albedo, forest_floor, conifer/grass
 * silhouettes, milky_way.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <tgmath.h>

#include "sg_oracle.h"

#ifndef REAL
#define REAL float
#endif
typedef REAL R;

#define MISS 0xffffffffu
#define LOCAL_BASE 0x80000000u
#define TERRAIN_ID 0x80100000u
#define ORC_STACK 40

static const R PI = 3.141592653589793;
static const R TAU = 6.283185307179586;
static const R METRES_PER_AU = 149597870700.0;
static const R STAR_DOT_PIXELS = 3.0;
#define ATMO_VIEW_STEPS 8u
#define ATMO_SUN_STEPS 4u
static const R POINT_DISC_RADIUS = 0.00136;
static const R RING_INDIRECT_CAP = 0.003;

typedef struct { float v[4]; } f4;
typedef struct {
    f4 cam_right, cam_up, cam_forward, viewport, atmo_center, atmo_rayleigh, atmo_params, ground_east, ground_up,
        ground_north;
    uint32_t ground_counts[4];
    uint32_t counts[4];
    uint32_t sampling[4];
    float background[4];
    uint32_t lights[4];
} Settings;
typedef struct {
    f4 center, material, low, ring_plane, ring_params, ring_color, axis_x, axis_y, surface_color, surface_params;
} Body;
typedef struct { f4 center, extent, albedo; } Prim;
typedef struct { f4 direction, radiance; } RayStar;

typedef struct { R x, y, z; } v3;
static inline v3 V(R x, R y, R z) { v3 r = {x, y, z}; return r; }
static inline v3 Vs(R s) { return V(s, s, s); }
static inline v3 V4(f4 a) { return V(a.v[0], a.v[1], a.v[2]); }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 dvv(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl(v3 a, R s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, R s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline R dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline R len(v3 a) { return sqrt(dot(a, a)); }
static inline v3 nrm(v3 a) { return dvs(a, len(a)); }
static inline R mn(R a, R b) { return a < b ? a : b; }
static inline R mx(R a, R b) { return a > b ? a : b; }
static inline R clampR(R x, R lo, R hi) { return mn(mx(x, lo), hi); }
static inline R mixR(R a, R b, R t) { return a * (1.0 - t) + b * t; }
static inline v3 mixV(v3 a, v3 b, R t) { return add(scl(a, 1.0 - t), scl(b, t)); }
static inline R smooth(R e0, R e1, R x) {
    R t = clampR((x - e0) / (e1 - e0), 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}
static inline R fract(R x) { return x - floor(x); }
static inline R maxc(v3 a) { return mx(a.x, mx(a.y, a.z)); }
static inline R lum(v3 c) { return dot(c, V(0.2126, 0.7152, 0.0722)); }
static inline v3 expV(v3 a) { return V(exp(a.x), exp(a.y), exp(a.z)); }
static inline R sel(R f, R t, int c) { return c ? t : f; }

typedef struct { v3 offset, direction; uint32_t anchor; } Ray;
typedef struct { R distance; uint32_t index; v3 normal; } Hit;
typedef struct { v3 direction; R pdf; } LightSample;
typedef struct { R distance; uint32_t index; v3 point; } RingHit;

typedef struct {
    const Settings *s;
    const Body *b;
    const Prim *objects;
    const float *sky;
    float *terrain;
    const RayStar *catalogue;
    uint32_t ncat;
    const uint32_t *cells;
    const uint32_t *milky;
    const float *ring_profile;
    uint32_t ring_profile_body;
    uint32_t view_steps, sun_steps; /* air march */
    R mie_height;                   /* 0: one scale height */
    /* per-dispatch counts.y (the sample index semantics of the host slices) */
    uint32_t accumulated;
} Ctx;

static Hit hit_miss(void) { Hit h = {1.0e30, MISS, {0, 0, 0}}; return h; }

/* Event counters for the algorithmic-work model (single-threaded runs only). */
enum {
    EV_PATHS, EV_SEGMENTS, EV_SPHERE, EV_SPHERE_HIT, EV_RING, EV_LIGHT_SAMPLE, EV_VISIBILITY, EV_RING_OPTICS,
    EV_ATMO_CELL, EV_DENSITY_STEP, EV_VNOISE, EV_BOX, EV_ROCK, EV_VEGETATION, EV_TERRAIN_NODE, EV_TRIANGLE,
    EV_VERTEX, EV_RNG, EV_COUNT
};
/* Counts are kept per stage: 0 for everything but the camera-air step, 1 for
 * the camera-air step (the view march of integrate_camera_air and
 * camera_air_transmittance), which the product runs as a kernel of its own. */
static uint64_t counters[2][EV_COUNT];
static _Thread_local uint64_t local_counters[2][EV_COUNT];
static _Thread_local int stage = 0;
static int counting = 0;
#define COUNT(e) do { if (counting) local_counters[stage][e] += 1; } while (0)
void orc_set_counting(int on) { counting = on; memset(counters, 0, sizeof(counters)); }
/* 2 x EV_COUNT values: all stages but the camera air, then the camera air. */
void orc_counters(uint64_t *out) { memcpy(out, counters, sizeof(counters)); }
/* Folds the calling thread's counts into the totals. */
static void flush_counters(void) {
    if (!counting) return;
#pragma omp critical(orc_counters)
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < EV_COUNT; ++i) {
            counters[k][i] += local_counters[k][i];
            local_counters[k][i] = 0;
        }
}
uint32_t orc_counter_count(void) { return EV_COUNT; }

/* Test knob (no product counterpart): with 0, planetshine in the air is never
 * thinned, the estimator the reference's energy test compares against
 * (src/motion_noise_tests.rs:354-360 sets the roulette minimum to 1). The
 * draw is still taken, so the paths' other draws do not shift. */
static int planetshine_thinning = 1;
void orc_set_planetshine_thinning(int on) { planetshine_thinning = on; }

/* --- RNG (src/shaders/pathtrace.wgsl:87-98) --- */
uint32_t orc_hash(uint32_t value) {
    uint32_t x = value;
    x = (x ^ (x >> 16)) * 0x7feb352du;
    x = (x ^ (x >> 15)) * 0x846ca68bu;
    return x ^ (x >> 16);
}
static R random_(uint32_t *state) {
    COUNT(EV_RNG);
    *state = *state * 747796405u + 2891336453u;
    uint32_t word = ((*state >> ((*state >> 28) + 4u)) ^ *state) * 277803737u;
    return (R)(((word >> 22) ^ word) >> 8) * (1.0 / 16777216.0);
}
double orc_random(uint32_t *state) { return (double)random_(state); }

static const Body *B(const Ctx *c, uint32_t i) { return &c->b[i]; }

/* src/shaders/pathtrace.wgsl:100-113 */
static v3 relative_center(const Ctx *c, Ray ray, uint32_t index) {
    if (ray.anchor == index) return neg(ray.offset);
    if (ray.anchor == MISS) return add(V4(B(c, index)->center), sub(V4(B(c, index)->low), ray.offset));
    return add(sub(V4(B(c, index)->center), V4(B(c, ray.anchor)->center)),
               sub(sub(V4(B(c, index)->low), V4(B(c, ray.anchor)->low)), ray.offset));
}

/* src/shaders/pathtrace.wgsl:134-171 */
static Hit sphere_hit_radius(const Ctx *c, Ray ray, uint32_t index, R radius) {
    COUNT(EV_SPHERE);
    v3 oc = relative_center(c, ray, index);
    R a = dot(ray.direction, ray.direction);
    R b = dot(oc, ray.direction);
    v3 pc = cross(oc, ray.direction);
    R h = a * radius * radius - dot(pc, pc);
    if (h < 0.0) return hit_miss();
    R root = sqrt(h);
    R q = b + sel(-root, root, b >= 0.0);
    if (q == 0.0) return hit_miss();
    R cc = dot(oc, oc) - radius * radius;
    R nearv = sel(cc / q, q / a, b < 0.0);
    R farv = sel(q / a, cc / q, b < 0.0);
    R distance = nearv;
    R delta = -root / a;
    if (distance <= 0.0) {
        distance = farv;
        delta = root / a;
    }
    if (distance <= 0.0) return hit_miss();
    COUNT(EV_SPHERE_HIT);
    v3 perpendicular = dvs(cross(ray.direction, pc), a);
    Hit r = {distance, index, nrm(add(neg(perpendicular), scl(ray.direction, delta)))};
    return r;
}
static Hit sphere_hit(const Ctx *c, Ray ray, uint32_t index) {
    return sphere_hit_radius(c, ray, index, B(c, index)->center.v[3]);
}

/* src/shaders/pathtrace.wgsl:187-194 */
static v3 basis_direction(v3 axis, R x, R y, R z) {
    v3 helper = V(0.0, 0.0, 1.0);
    if (fabs(axis.z) > 0.9) helper = V(0.0, 1.0, 0.0);
    v3 tangent = nrm(cross(helper, axis));
    return nrm(add(add(scl(tangent, x), scl(cross(axis, tangent), y)), scl(axis, z)));
}
/* src/shaders/pathtrace.wgsl:197-211 */
static R cone_width(const Ctx *c, Ray ray, uint32_t index) {
    v3 oc = relative_center(c, ray, index);
    R r = B(c, index)->center.v[3];
    R sin2 = clampR(r * r / dot(oc, oc), 1.0e-20, 1.0);
    return sin2 / (1.0 + sqrt(1.0 - sin2));
}
static uint32_t light_count(const Ctx *c) { return c->s->lights[3] <= 3u ? c->s->lights[3] : c->s->counts[0]; }
static uint32_t light_index(const Ctx *c, uint32_t slot) { return c->s->lights[3] <= 3u ? c->s->lights[slot] : slot; }
/* src/shaders/pathtrace.wgsl:214-234 */
static LightSample sample_light_stratum(const Ctx *c, Ray ray, uint32_t index, uint32_t stratum, uint32_t count,
                                        uint32_t *rng) {
    COUNT(EV_LIGHT_SAMPLE);
    v3 axis = nrm(relative_center(c, ray, index));
    R width = cone_width(c, ray, index);
    R radial = random_(rng);
    R azimuth = random_(rng);
    if (count == 4u) {
        radial = ((R)(stratum & 1u) + radial) * 0.5;
        azimuth = ((R)(stratum >> 1u) + azimuth) * 0.5;
    }
    R one_minus_cos = radial * width;
    R sin_theta = sqrt(one_minus_cos * (2.0 - one_minus_cos));
    R phi = TAU * azimuth;
    LightSample l = {basis_direction(axis, sin_theta * cos(phi), sin_theta * sin(phi), 1.0 - one_minus_cos),
                     1.0 / (TAU * width)};
    return l;
}
static LightSample sample_light(const Ctx *c, Ray ray, uint32_t index, uint32_t *rng) {
    return sample_light_stratum(c, ray, index, 0u, 1u, rng);
}
/* src/shaders/pathtrace.wgsl:239-263 */
static uint32_t eclipse_light_samples(const Ctx *c, Ray ray, uint32_t index) {
    if (c->accumulated != 0u) return 1u;
    v3 source = relative_center(c, ray, index);
    R distance = len(source);
    v3 axis = dvs(source, distance);
    R source_chord = sqrt(2.0 * cone_width(c, ray, index));
    for (uint32_t b = 0; b < c->s->counts[0]; ++b) {
        if (b == index) continue;
        v3 offset = relative_center(c, ray, b);
        R d2 = dot(offset, offset);
        R radius = B(c, b)->center.v[3];
        if (d2 <= radius * radius) continue;
        R d = sqrt(d2);
        if (d - radius >= distance + B(c, index)->center.v[3]) continue;
        R sin2 = clampR(radius * radius / d2, 0.0, 1.0);
        R limb_chord = sqrt(2.0 * sin2 / (1.0 + sqrt(1.0 - sin2)));
        R separation = len(sub(axis, dvs(offset, d)));
        if (separation < source_chord + limb_chord && separation + source_chord > limb_chord) return 4u;
    }
    return 1u;
}
static R power_weight(R a, R b) {
    R ratio = b / mx(a, 1.0e-30);
    return 1.0 / (1.0 + ratio * ratio);
}

/* --- background (src/shaders/pathtrace.wgsl:115-128, 269-334) --- */
static v3 world_direction(const Ctx *c, v3 d) {
    const Settings *s = c->s;
    return nrm(sub(add(scl(V4(s->cam_right), d.x), scl(V4(s->cam_up), d.y)), scl(V4(s->cam_forward), d.z)));
}
static int inside_star_dot(const Ctx *c, v3 view_direction, v3 star_direction) {
    const Settings *s = c->s;
    v3 sv = V(dot(star_direction, V4(s->cam_right)), dot(star_direction, V4(s->cam_up)),
              -dot(star_direction, V4(s->cam_forward)));
    if (sv.z >= 0.0) return 0;
    R dx = view_direction.x / -view_direction.z - sv.x / -sv.z;
    R dy = view_direction.y / -view_direction.z - sv.y / -sv.z;
    R pixel_scale = (R)s->viewport.v[1] / (2.0 * (R)s->cam_forward.v[3]);
    dx *= pixel_scale;
    dy *= pixel_scale;
    return dx * dx + dy * dy <= STAR_DOT_PIXELS * STAR_DOT_PIXELS;
}
static v3 environment(const Ctx *c, v3 view_direction, int primary) {
    v3 light = Vs(0.0);
    if (!primary || c->ncat == 0) return light;
    v3 direction = world_direction(c, view_direction);
    R phi = atan2(direction.y, direction.x) + PI;
    R theta = acos(clampR(direction.z, -1.0, 1.0));
    uint32_t x = (uint32_t)(phi * (256.0 / TAU));
    uint32_t y = (uint32_t)(theta * (128.0 / PI));
    if (x > 255u) x = 255u;
    if (y > 127u) y = 127u;
    uint32_t cell = y * 256u + x;
    for (uint32_t j = c->cells[cell]; j < c->cells[cell + 1u]; ++j) {
        const RayStar *star = &c->catalogue[c->cells[32769u + j]];
        v3 sd = V4(star->direction);
        v3 perpendicular = cross(direction, sd);
        if (dot(direction, sd) > 0.0 && dot(perpendicular, perpendicular) <= (R)star->direction.v[3]) {
            if (inside_star_dot(c, view_direction, sd)) light = add(light, V4(star->radiance));
        }
    }
    return light;
}
static v3 milky_way_texel(const Ctx *c, int px, int py) {
    int width = (int)c->s->background[1];
    int height = (int)c->s->background[2];
    int x = (px + width) % width;
    int y = py < 0 ? 0 : (py > height - 1 ? height - 1 : py);
    uint32_t packed = c->milky[(uint32_t)(y * width + x)];
    return dvs(V((R)(packed & 1023u), (R)((packed >> 10) & 1023u), (R)((packed >> 20) & 1023u)), 1023.0);
}
static v3 milky_way(const Ctx *c, v3 view_direction) {
    if (c->s->background[0] == 0.0f || !c->milky) return Vs(0.0);
    v3 direction = world_direction(c, view_direction);
    v3 pole = nrm(V(0.35, 0.82, 0.45));
    R a = -1.0 / (1.0 + pole.z);
    v3 right = V(pole.x * pole.y * a, 1.0 + pole.y * pole.y * a, -pole.y);
    v3 forward = cross(pole, right);
    R longitude = atan2(dot(direction, forward), dot(direction, right));
    R u = 0.5 - longitude / TAU;
    R v = acos(clampR(dot(direction, pole), -1.0, 1.0)) / PI;
    R px = u * (R)c->s->background[1] - 0.5, py = v * (R)c->s->background[2] - 0.5;
    int lx = (int)floor(px), ly = (int)floor(py);
    R wx = fract(px), wy = fract(py);
    v3 l = mixV(mixV(milky_way_texel(c, lx, ly), milky_way_texel(c, lx + 1, ly), wx),
                mixV(milky_way_texel(c, lx, ly + 1), milky_way_texel(c, lx + 1, ly + 1), wx), wy);
    return scl(l, (R)c->s->background[0]);
}

/* --- procedural albedo (src/shaders/pathtrace.wgsl:338-470, 897-925) --- */
static R hash31(v3 p) {
    v3 q = V(fract(p.x * 0.1031), fract(p.y * 0.1030), fract(p.z * 0.0973));
    /* the dot product as two fused multiply-adds, in the order the HIP
     * build evaluates it (WGSL leaves the contraction to the implementation) */
    R d = fma(q.x, q.y + 33.33, fma(q.y, q.x + 33.33, q.z * (q.z + 33.33)));
    q = add(q, Vs(d));
    return fract((q.x + q.y) * q.z);
}
static R vnoise(v3 p) {
    COUNT(EV_VNOISE);
    v3 i = V(floor(p.x), floor(p.y), floor(p.z));
    v3 f = sub(p, i);
    v3 u = mul(mul(f, f), sub(Vs(3.0), scl(f, 2.0)));
    R a = mixR(hash31(i), hash31(add(i, V(1, 0, 0))), u.x);
    R b = mixR(hash31(add(i, V(0, 1, 0))), hash31(add(i, V(1, 1, 0))), u.x);
    R cc = mixR(hash31(add(i, V(0, 0, 1))), hash31(add(i, V(1, 0, 1))), u.x);
    R d = mixR(hash31(add(i, V(0, 1, 1))), hash31(add(i, V(1, 1, 1))), u.x);
    return mixR(mixR(a, b, u.y), mixR(cc, d, u.y), u.z);
}
static R fbm(v3 p, uint32_t octaves) {
    R amplitude = 0.5, sum = 0.0, norm = 0.0;
    for (uint32_t o = 0; o < octaves; ++o) {
        sum += amplitude * vnoise(p);
        norm += amplitude;
        /* mat3x3 columns (0,1.6,1.2), (-1.6,0.72,-0.96), (-1.2,-0.96,1.28) times p */
        v3 r = add(add(scl(V(0.00, 1.60, 1.20), p.x), scl(V(-1.60, 0.72, -0.96), p.y)),
                   scl(V(-1.20, -0.96, 1.28), p.z));
        p = add(r, V(17.3, 5.1, 11.7));
        amplitude *= 0.5;
    }
    return sum / norm;
}
static v3 body_local(const Body *b, v3 n) {
    v3 x = V4(b->axis_x), y = V4(b->axis_y);
    return V(dot(n, x), dot(n, y), dot(n, cross(x, y)));
}
static v3 seed_offset(R seed) {
    return scl(V(hash31(V(seed, 1.0, 7.0)), hash31(V(3.0, seed, 5.0)), hash31(V(9.0, 2.0, seed))), 512.0);
}
static v3 gas_giant_albedo(const Body *body, v3 local) {
    v3 off = seed_offset(body->axis_y.v[3]);
    v3 zone = V4(body->material), belt = V4(body->surface_color);
    R turbulence = body->surface_color.v[3];
    R bands = body->surface_params.v[0];
    uint32_t storms = (uint32_t)((R)body->surface_params.v[1] + 0.5);
    v3 q = local;
    R masks[3] = {0, 0, 0}, rims[3] = {0, 0, 0};
    for (uint32_t k = 0; k < storms && k < 3u; ++k) {
        R fk = (R)k;
        R lat = (hash31(add(off, V(7.1 * fk, 1.3, 2.9))) - 0.5) * 1.1;
        R lon = hash31(add(off, V(3.7, 5.3 * fk, 8.1))) * TAU;
        R size = sel(0.05, 0.11, k == 0u) * (0.8 + 0.4 * hash31(add(off, V(fk, 9.0, 4.0))));
        v3 centre = V(cos(lat) * cos(lon), cos(lat) * sin(lon), sin(lat));
        if (dot(q, centre) <= 0.0) continue;
        v3 east = V(-sin(lon), cos(lon), 0.0);
        v3 north = cross(centre, east);
        R dx = dot(sub(q, centre), east) / (2.0 * size);
        R dy = dot(sub(q, centre), north) / size;
        R e = sqrt(dx * dx + dy * dy);
        R influence = 1.0 - smooth(0.0, 2.4, e);
        R angle = sel(-3.2, 3.2, lat < 0.0) * influence * influence;
        R cs = cos(angle), sn = sin(angle);
        R rx = cs * dx - sn * dy;
        R ry = sn * dx + cs * dy;
        q = nrm(add(add(q, scl(east, (rx - dx) * 2.0 * size)), scl(north, (ry - dy) * size)));
        R ragged = e + 0.25 * (fbm(add(scl(q, 18.0), off), 3) - 0.5);
        masks[k] = 1.0 - smooth(0.7, 1.0, ragged);
        rims[k] = smooth(0.75, 1.0, ragged) * (1.0 - smooth(1.0, 1.35, ragged));
    }
    v3 p = add(V(q.x * 2.2, q.y * 2.2, q.z * 9.0), off);
    R warp1 = fbm(p, 5) - 0.5;
    R warp2 = fbm(add(add(scl(p, 2.1), V(5.2, 1.3, 7.7)), Vs(4.0 * warp1)), 5) - 0.5;
    R lat = q.z + 0.07 * sin(3.0 * PI * q.z + off.z) + 0.04 * sin(5.0 * PI * q.z + off.y);
    R shear = 0.35 + 0.65 * fabs(cos(PI * bands * lat + off.x));
    R lon = atan2(q.y, q.x);
    R ripple = sin(lon * (9.0 + floor(bands)) + 6.0 * warp1 + off.y) * 0.5 + sin(lon * 23.0 + 9.0 * warp2 + off.z) * 0.25;
    R mu = lat + turbulence * shear * (0.16 * warp1 + 0.10 * warp2 + 0.025 * ripple);
    R wave = 0.5 + 0.5 * sin(PI * bands * mu + off.x);
    R irregular = fbm(V(mu * bands * 1.7, off.y, off.z), 3);
    R profile = 0.5 * wave + 0.5 * irregular;
    R tone = fbm(V(mu * bands * 0.9, off.z, off.x), 2);
    R fine = fbm(add(add(V(q.x * 7.0, q.y * 7.0, q.z * 70.0), off), Vs(3.0 * warp2)), 4);
    R eddies = fbm(add(scl(p, 5.0), scl(V(warp1, warp2, warp1), 6.0)), 4);
    v3 color = mixV(belt, zone, smooth(0.3, 0.75, profile));
    color = mixV(color, mul(color, add(Vs(0.55), scl(color, 0.5))), smooth(0.35, 0.75, tone) * 0.8);
    color = scl(color, (0.78 + 0.44 * fine) * (0.9 + 0.2 * eddies));
    R polar = smooth(0.55, 0.92, fabs(local.z));
    color = mixV(color, scl(mixV(zone, belt, 0.55), 0.8 + 0.2 * fine), polar * 0.8);
    for (uint32_t k = 0; k < storms && k < 3u; ++k) {
        v3 tint = k == 0u ? scl(mul(belt, belt), 1.5) : scl(zone, 1.15);
        R core = 1.0 - 0.25 * (1.0 - smooth(0.0, 0.6, masks[k] - rims[k]));
        color = mixV(color, scl(scl(tint, 0.7 + 0.6 * fine), core), masks[k] * 0.92);
        color = mixV(color, scl(zone, 1.1), rims[k] * 0.45);
    }
    return color;
}
/* --- solid-body surfaces (src/shaders/pathtrace.wgsl:472-895) --- */
static R surface_gaussian(R d) { return exp(-(d * d)); } /* wgsl:477 */
static v3 hash33(v3 p) { /* wgsl:481 */
    return V(hash31(p), hash31(add(p, V(19.1, 7.3, 3.7))), hash31(add(p, V(5.9, 31.7, 11.3))));
}
static R maria(v3 q, v3 off) { /* wgsl:492 */
    R warp = fbm(add(scl(q, 2.0), off), 4) - 0.5;
    return smooth(0.5, 0.56, fbm(add(add(scl(q, 1.2), off), Vs(1.5 * warp)), 5));
}
/* wgsl:502-511: the crater centre owned by a cell, or found = 0. */
static int cell_centre(v3 cell, v3 key, R scale, v3 off, v3 *centre) {
    for (uint32_t t = 0; t < 3u; ++t) {
        v3 c = nrm(sub(add(cell, hash33(add(key, Vs((R)t * 17.0)))), off));
        v3 g = add(scl(c, scale), off);
        if (floor(g.x) == cell.x && floor(g.y) == cell.y && floor(g.z) == cell.z) {
            *centre = c;
            return 1;
        }
    }
    return 0;
}
static v3 crater_grid(v3 base, uint32_t i) {
    return add(base, V((R)(i & 1u), (R)((i >> 1) & 1u), (R)((i >> 2) & 1u)));
}
/* wgsl:527-633: one octave; st[0] floor multiplier, st[1] relief. */
static void crater_layer(v3 q, R scale, R density, v3 off, v3 body_off, R flood, R contrast, R rays, R st[2]) {
    v3 p = add(scl(q, scale), off);
    v3 base = V(floor(p.x - 0.5), floor(p.y - 0.5), floor(p.z - 0.5));
    uint32_t slots[6];
    R ages[6];
    v3 centres[6];
    uint32_t count = 0;
    for (uint32_t i = 0; i < 16u; ++i) {
        v3 grid = crater_grid(base, i);
        v3 gap = V(mx(mx(grid.x - p.x, p.x - grid.x - 1.0), 0.0), mx(mx(grid.y - p.y, p.y - grid.y - 1.0), 0.0),
                   mx(mx(grid.z - p.z, p.z - grid.z - 1.0), 0.0));
        if (dot(gap, gap) > 0.2343) continue;
        v3 key = add(grid, Vs((R)(i >> 3) * 1000.0));
        if (hash31(add(key, V(71.3, 3.1, 9.7))) > density) continue;
        v3 centre;
        if (!cell_centre(grid, key, scale, off, &centre)) continue;
        R size = hash31(add(key, V(3.1, 8.3, 1.7)));
        R radius = (0.06 + 0.16 * size * size) / scale;
        if (len(sub(q, centre)) > 2.2 * radius) continue;
        R age = hash31(add(key, V(29.0, 61.0, 7.0)));
        if (flood > 0.0 && age < 0.7 && maria(centre, body_off) > 0.5) continue;
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
        for (; j > 0u && ages[j - 1] > age; --j) {
            ages[j] = ages[j - 1];
            slots[j] = slots[j - 1];
            centres[j] = centres[j - 1];
        }
        ages[j] = age;
        slots[j] = i;
        centres[j] = centre;
        ++count;
    }
    R floor_shade = 1.0, relief = st[1];
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t i = slots[k];
        v3 cell = add(crater_grid(base, i), Vs((R)(i >> 3) * 1000.0));
        v3 centre = centres[k];
        R size = hash31(add(cell, V(3.1, 8.3, 1.7)));
        R radius = (0.06 + 0.16 * size * size) / scale;
        v3 offset = sub(q, centre);
        R d = len(offset) / radius;
        v3 east = cross(V(0.0, 0.0, 1.0), centre);
        east = dvs(east, mx(len(east), 1.0e-4));
        v3 north = cross(centre, east);
        R angle = atan2(dot(offset, north), dot(offset, east));
        v3 phase = scl(hash33(add(cell, V(41.0, 5.0, 23.0))), TAU);
        R wobble = 0.025 * (sin(3.0 * angle + phase.x) + sin(5.0 * angle + phase.y) + 0.6 * sin(8.0 * angle + phase.z));
        R r = d + wobble;
        floor_shade = mn(floor_shade, 1.0 - 0.18 * contrast * (1.0 - smooth(0.55, 0.9, r)));
        R rim = 0.22 * contrast * surface_gaussian((r - 1.0) / 0.13);
        R footprint = 1.0 - smooth(1.1, 1.3, r);
        relief = mixR(relief, rim, footprint);
        R blanket = (1.0 - smooth(1.0, 1.9, r)) * (1.0 - footprint);
        relief = mixR(relief, 0.0, blanket * 0.3) + 0.1 * contrast * blanket;
        R fresh = hash31(add(cell, V(13.0, 2.0, 5.0)));
        if (rays > 0.0 && fresh > 0.82) {
            R streaks = pow(vnoise(V(cos(angle) * 9.0, sin(angle) * 9.0, fresh * 40.0)), 3.0);
            relief += contrast * (rays * 0.9 * streaks * (1.0 - smooth(1.0, 2.2, d)) * (d >= 0.9 ? 1.0 : 0.0) +
                                  rays * 0.25 * (1.0 - smooth(0.0, 1.0, d)));
        }
    }
    st[0] *= floor_shade;
    st[1] = relief;
}
static R craters(v3 q, R density, v3 off, R flood, R contrast, R rays) { /* wgsl:636 */
    if (density <= 0.0) return 1.0;
    R st[2] = {1.0, 0.0};
    crater_layer(q, 1.5, density, off, off, flood, contrast, rays, st);
    crater_layer(q, 3.5, density, add(off, Vs(101.0)), off, flood, contrast, rays * 0.8, st);
    crater_layer(q, 8.0, density, add(off, Vs(203.0)), off, flood, contrast * 0.85, rays * 0.5, st);
    crater_layer(q, 18.0, density, add(off, Vs(307.0)), off, 0.0, contrast * 0.7, 0.0, st);
    crater_layer(q, 40.0, density, add(off, Vs(401.0)), off, 0.0, contrast * 0.55, 0.0, st);
    return st[0] * (1.0 + st[1]);
}
static R polar_caps(v3 q, R extent, v3 off) { /* wgsl:650 */
    if (extent <= 0.0) return 0.0;
    R edge = 1.0 - extent + 0.12 * (fbm(add(scl(q, 5.0), off), 5) - 0.5) + 0.05 * (fbm(add(scl(q, 18.0), off), 3) - 0.5);
    return smooth(edge - 0.015, edge + 0.015, fabs(q.z));
}
static v3 cratered_albedo(const Body *body, v3 q) { /* wgsl:660 */
    v3 off = seed_offset(body->axis_y.v[3]);
    R mare = maria(q, off);
    R grain = fbm(add(scl(q, 45.0), off), 3);
    v3 color = mixV(V4(body->material), V4(body->surface_color), mare);
    color = scl(color, 0.9 + 0.2 * grain);
    return scl(color, craters(q, body->surface_color.v[3], off, 1.0, 1.0, 1.0));
}
static v3 rocky_albedo(const Body *body, v3 q) { /* wgsl:673 */
    v3 off = seed_offset(body->axis_y.v[3]);
    R wx = fbm(add(scl(q, 1.8), off), 4) - 0.5;
    R wy = fbm(add(add(scl(q, 1.8), off), Vs(7.0)), 4) - 0.5;
    R dark = smooth(0.5, 0.64, fbm(add(add(scl(q, 2.2), off), V(2.0 * wx, 2.0 * wy, wx)), 6));
    R tone = fbm(add(add(scl(q, 5.0), off), Vs(3.0)), 4);
    v3 color = mixV(V4(body->material), V4(body->surface_color), dark * 0.85);
    color = scl(color, 0.85 + 0.3 * tone);
    R streak = fbm(add(add(V(q.x * 4.0, q.y * 4.0, q.z * 18.0), off), Vs(9.0)), 3);
    color = scl(color, 0.95 + 0.1 * streak);
    color = scl(color, craters(q, body->surface_color.v[3], off, 0.0, 0.45, 0.0));
    R ice = polar_caps(q, body->surface_params.v[0], off);
    return mixV(color, V(0.92, 0.92, 0.94), ice);
}
static R ridge(R n, R width) { return 1.0 - smooth(0.0, width, fabs(n - 0.5)); } /* wgsl:487 */
static v3 ice_plates(v3 p) { /* wgsl:691-709 */
    v3 base = V(floor(p.x), floor(p.y), floor(p.z));
    R nearest = 10.0, second = 10.0, tone = 0.0;
    for (int i = 0; i < 27; ++i) {
        v3 cell = add(base, V((R)(i % 3 - 1), (R)((i / 3) % 3 - 1), (R)(i / 9 - 1)));
        v3 delta = sub(add(add(cell, Vs(0.2)), scl(hash33(cell), 0.6)), p);
        R d = dot(delta, delta);
        if (d < nearest) {
            second = nearest;
            nearest = d;
            tone = hash31(add(cell, Vs(57.0)));
        } else {
            second = mn(second, d);
        }
    }
    return V(sqrt(nearest), sqrt(second) - sqrt(nearest), tone);
}
static v3 icy_albedo(const Body *body, v3 q) { /* wgsl:714-767 */
    v3 off = seed_offset(body->axis_y.v[3]);
    v3 ice = V4(body->material);
    v3 rust = mul(V4(body->surface_color), V(0.72, 0.62, 0.56));
    v3 warp = sub(V(fbm(add(add(scl(q, 3.1), off), Vs(11.0)), 3), fbm(add(add(scl(q, 3.1), off), Vs(23.0)), 3),
                    fbm(add(add(scl(q, 3.1), off), Vs(37.0)), 3)),
                  Vs(0.5));
    R terrain = fbm(add(add(scl(q, 4.2), off), warp), 5);
    R rubble = fbm(add(add(scl(q, 85.0), off), scl(warp, 7.0)), 4);
    R chaos = smooth(0.42, 0.65, terrain + 0.14 * (rubble - 0.5));
    v3 plates = ice_plates(add(add(scl(q, 72.0), off), scl(warp, 4.0)));
    R seams = 1.0 - smooth(0.015, 0.07, plates.y);
    R dust = chaos * (0.2 + 0.46 * rubble + 0.18 * plates.z);
    v3 color = mixV(ice, rust, dust);
    color = mixV(color, scl(rust, 0.8), seams * chaos * 0.35);
    R ridged_plain = ridge(fbm(add(add(scl(q, 34.0), off), Vs(83.0)), 3), 0.025);
    color = mixV(color, rust, ridged_plain * (0.1 + 0.14 * chaos));
    color = scl(color, 0.87 + 0.22 * fbm(add(scl(q, 180.0), off), 3));
    v3 warped = add(q, scl(warp, 0.1));
    R roughness = 0.0015 * (rubble - 0.5);
    R shoulder = 0.0, fissure = 0.0, fresh_ridge = 0.0;
    for (uint32_t i = 0; i < 64u; ++i) {
        v3 h = hash33(add(off, V((R)i * 3.7, 1.0, 2.0)));
        v3 family = hash33(add(off, V((R)(i / 8u) * 17.3, 21.0, 7.0)));
        v3 n = nrm(add(sub(scl(family, 2.0), Vs(1.0)), scl(sub(scl(h, 2.0), Vs(1.0)), 0.28)));
        R shift = (hash31(add(off, V((R)i, 9.0, 4.0))) - 0.5) * 1.4;
        int old_band = i < 10u;
        R width = (old_band ? 0.0015 + 0.0022 * h.z : 0.0005 + 0.0011 * h.z) *
                  (0.65 + 0.8 * vnoise(add(scl(q, 17.0), scl(h, 29.0))));
        R bend = 0.03 * (vnoise(add(scl(q, 6.0), scl(family, 31.0))) - 0.5) +
                 0.008 * (vnoise(add(scl(q, 24.0), scl(family, 37.0))) - 0.5);
        R dist = fabs(dot(warped, n) - shift + roughness + bend);
        R along = smooth(0.28, 0.62, vnoise(add(scl(q, 3.5), scl(h, 50.0))));
        shoulder = mx(shoulder, (1.0 - smooth(width, width * 4.0, dist)) * along);
        R double_ridge = fabs(dist - width * 0.85);
        R single = 1.0 - smooth(width * 0.25, width, dist);
        R paired = 1.0 - smooth(width * 0.15, width * 0.55, double_ridge);
        fissure = mx(fissure, (old_band ? paired : single) * along);
        fresh_ridge = mx(fresh_ridge, (1.0 - smooth(width * 0.15, width * 0.35, fabs(dist - width * 1.8))) * along);
    }
    color = mixV(color, rust, shoulder * 0.17);
    color = mixV(color, scl(rust, 0.7), fissure * 0.8);
    color = mixV(color, ice, fresh_ridge * 0.13);
    return scl(color, craters(q, body->surface_color.v[3] * 0.2, off, 0.0, 0.3, 0.4));
}
static v3 volcanic_albedo(const Body *body, v3 q) { /* wgsl:770-851 */
    v3 off = seed_offset(body->axis_y.v[3]);
    v3 sulphur = V4(body->material), deposit = V4(body->surface_color);
    v3 warp = sub(V(fbm(add(scl(q, 4.7), off), 3), fbm(add(add(scl(q, 4.7), off), Vs(19.0)), 3),
                    fbm(add(add(scl(q, 4.7), off), Vs(41.0)), 3)),
                  Vs(0.5));
    R terrain = fbm(add(add(scl(q, 10.0), off), scl(warp, 2.0)), 6);
    R weathered = smooth(0.38, 0.65, fbm(add(add(scl(q, 3.7), off), Vs(11.0)), 5));
    R fine = fbm(add(add(scl(q, 90.0), off), scl(warp, 3.0)), 4);
    v3 crust = mixV(mul(sulphur, V(0.43, 0.51, 0.6)), deposit, 0.12);
    v3 color = scl(sulphur, 0.7 + 0.5 * terrain);
    color = mixV(color, crust, weathered * (0.45 + 0.5 * terrain));
    R frost = smooth(0.38, 0.74, fbm(add(add(add(scl(q, 6.0), off), Vs(29.0)), warp), 5) + 0.2 * (fine - 0.5));
    v3 frost_color = mixV(sulphur, V(0.72, 0.7, 0.62), 0.8);
    color = mixV(color, frost_color, frost * 0.7);
    R flow_edges = ridge(fbm(add(add(scl(q, 35.0), off), scl(warp, 2.0)), 4), 0.025);
    color = mixV(color, crust, flow_edges * weathered * 0.4);
    color = scl(color, (0.82 + 0.32 * fine) * (0.94 + 0.13 * fbm(add(scl(q, 210.0), off), 3)));
    R dark = 0.0, ash = 0.0, red = 0.0, white = 0.0;
    R density = body->surface_color.v[3];
    R scale = 5.5;
    v3 p = add(add(scl(q, scale), off), Vs(300.0));
    v3 base = V(floor(p.x), floor(p.y), floor(p.z));
    for (int i = 0; i < 27; ++i) {
        v3 cell = add(base, V((R)(i % 3 - 1), (R)((i / 3) % 3 - 1), (R)(i / 9 - 1)));
        if (hash31(add(cell, V(4.0, 41.0, 9.0))) > 0.6 * density) continue;
        v3 centre;
        if (!cell_centre(cell, cell, scale, add(off, Vs(300.0)), &centre)) continue;
        R size = hash31(add(cell, V(8.0, 1.0, 3.0)));
        R radius = (0.025 + 0.11 * size * size * size) / scale;
        v3 offset = sub(q, centre);
        R reach = len(offset) * scale;
        if (reach > 0.95) continue;
        R cutoff = 1.0 - smooth(0.75, 0.95, reach);
        v3 tangent = cross(centre, V(0.37, 0.83, 0.41));
        v3 east = dvs(tangent, mx(len(tangent), 1.0e-5));
        v3 north = cross(centre, east);
        R angle = TAU * hash31(add(cell, Vs(91.0)));
        R ux = dot(offset, east) / radius, uy = dot(offset, north) / radius;
        R ca = cos(angle), sa = sin(angle);
        R vx = ux * ca - uy * sa, vy = ux * sa + uy * ca;
        R ragged = fbm(add(V(vx * 2.7, vy * 2.7, 0.0), scl(cell, 7.0)), 3) - 0.5;
        R ex = vx * (1.0 / (0.65 + 0.9 * size));
        R d = sqrt(ex * ex + vy * vy) + 0.42 * ragged;
        R pit = 1.0 - smooth(0.68, 0.86, d);
        R age = hash31(add(cell, Vs(31.0)));
        R fx = vx - (0.2 + 2.3 * age), fy = vy - 0.35 * sin(vx * 1.3 + angle);
        R sx = fx * (0.75 - 0.48 * age), sy = fy * (0.85 + 0.35 * size);
        R flow_shape = sqrt(sx * sx + sy * sy) + 0.65 * ragged +
                       0.35 * (fbm(add(V(vx * 1.8, vy * 1.8, 0.0), scl(cell, 3.0)), 4) - 0.5);
        R flow = (1.0 - smooth(0.72, 0.9, flow_shape)) * smooth(0.35, 0.85, age) * 0.85;
        dark = mx(dark, mx(pit, flow) * cutoff);
        R plume = hash31(add(cell, V(2.0, 6.0, 17.0)));
        R hx = (vx - 0.5) * 0.16, hy = (vy - 0.3) * 0.23;
        R halo = exp(-(hx * hx + hy * hy)) * cutoff;
        ash = mx(ash, halo * (0.25 + 0.35 * age));
        white = mx(white, halo * (plume >= 0.45 ? 1.0 : 0.0) * (1.0 - (plume >= 0.83 ? 1.0 : 0.0)));
        R apron = surface_gaussian((sqrt(vx * vx + vy * vy) + ragged - 3.5) / 1.6) *
                  (0.25 + 0.75 * smooth(-1.5, 2.0, vy));
        red = mx(red, apron * (plume >= 0.83 ? 1.0 : 0.0) * cutoff);
    }
    color = mixV(color, scl(crust, 0.7), ash * 0.45);
    color = mixV(color, frost_color, white * 0.4);
    color = mixV(color, mul(deposit, V(0.85, 0.7, 0.65)), red * 0.6);
    return mixV(color, V(0.035, 0.032, 0.026), dark * 0.96);
}
static v3 terran_albedo(const Body *body, v3 q) { /* wgsl:854-895 */
    v3 off = seed_offset(body->axis_y.v[3]);
    v3 land_base = V4(body->material), ocean_color = V4(body->surface_color);
    v3 warp = sub(V(fbm(add(scl(q, 1.5), off), 4), fbm(add(add(scl(q, 1.5), off), Vs(13.0)), 4),
                    fbm(add(add(scl(q, 1.5), off), Vs(29.0)), 4)),
                  Vs(0.5));
    R height = fbm(add(add(scl(q, 1.1), off), scl(warp, 0.9)), 7);
    R sea = 0.5 + ((R)body->surface_params.v[1] - 0.5) * 0.3;
    R coast = smooth(sea, sea + 0.004, height);
    R shallow = 1.0 - smooth(sea - 0.04, sea, height);
    v3 water = mixV(add(scl(ocean_color, 1.8), V(0.0, 0.04, 0.05)), ocean_color, shallow);
    R moisture = fbm(add(add(scl(q, 3.0), off), Vs(41.0)), 5) + 0.25 * (fabs(q.z) - 0.35);
    v3 green = mul(land_base, V(0.45, 0.62, 0.38));
    v3 land = mixV(land_base, green, smooth(0.45, 0.6, moisture));
    R altitude = smooth(sea + 0.08, sea + 0.2, height);
    land = mixV(land, V(0.42, 0.38, 0.34), altitude * 0.6);
    land = mixV(land, V(0.9, 0.9, 0.92), smooth(sea + 0.19, sea + 0.24, height));
    land = scl(land, 0.88 + 0.24 * fbm(add(scl(q, 25.0), off), 3));
    v3 color = mixV(water, land, coast);
    R ice = polar_caps(q, body->surface_params.v[0], off);
    color = mixV(color, V(0.9, 0.92, 0.95), ice);
    R cover = body->surface_params.v[2];
    if (cover > 0.0) {
        v3 w1 = sub(V(fbm(add(add(scl(q, 1.6), off), Vs(77.0)), 4), fbm(add(add(scl(q, 1.6), off), Vs(91.0)), 4), 0.5),
                    Vs(0.5));
        v3 w2 = sub(V(fbm(add(add(scl(q, 4.0), off), scl(w1, 3.0)), 4),
                      fbm(add(add(add(scl(q, 4.0), off), Vs(17.0)), scl(w1, 3.0)), 4), 0.5),
                    Vs(0.5));
        v3 cq = add(add(add(V(q.x * 2.5, q.y * 2.5, q.z * 7.0), off), scl(w1, 2.5)), scl(w2, 1.2));
        R cloud = fbm(cq, 7);
        R lat = fabs(q.z);
        R belts = 0.06 * (surface_gaussian(lat / 0.12) + surface_gaussian((lat - 0.72) / 0.15)) -
                  0.04 * surface_gaussian((lat - 0.42) / 0.1);
        R threshold = 0.62 - 0.22 * cover - belts;
        R mask = smooth(threshold - 0.04, threshold + 0.14, cloud);
        color = mixV(color, V(0.93, 0.94, 0.96), mask * 0.95);
    }
    return color;
}
/* Unclamped albedo of the procedural styles 1-6 at the local point q (codes
 * above 6 are terran, as in the WGSL switch). */
static v3 style_albedo(const Body *body, v3 q) {
    uint32_t style = (uint32_t)((R)body->axis_x.v[3] + 0.5);
    if (style == 1u) return gas_giant_albedo(body, q);
    if (style == 2u) return cratered_albedo(body, q);
    if (style == 3u) return rocky_albedo(body, q);
    if (style == 4u) return icy_albedo(body, q);
    if (style == 5u) return volcanic_albedo(body, q);
    return terran_albedo(body, q);
}
static v3 clamp01(v3 a) { return V(clampR(a.x, 0.0, 1.0), clampR(a.y, 0.0, 1.0), clampR(a.z, 0.0, 1.0)); }
static v3 surface_albedo(const Body *body, v3 normal) {
    v3 local = body_local(body, normal);
    uint32_t style = (uint32_t)((R)body->axis_x.v[3] + 0.5);
    if (style > 0u) return clamp01(style_albedo(body, local));
    R mottle = 1.0;
    if (body->low.v[3] > 0.0f) {
        v3 p = add(scl(local, 2.4), seed_offset(body->axis_y.v[3]));
        R amplitude = 0.5, noise = 0.0;
        for (uint32_t o = 0; o < 5u; ++o) {
            noise += amplitude * vnoise(p);
            p = scl(p, 2.07);
            amplitude *= 0.5;
        }
        mottle = mixR(1.0, 0.68 + 0.64 * noise, body->low.v[3]);
    }
    return clamp01(scl(V4(body->material), mottle));
}

/* --- local frame (src/shaders/ground.wgsl:17-61) --- */
static v3 camera_origin(const Ctx *c, Ray ray) {
    if (ray.anchor == MISS) return ray.offset;
    return add(V4(B(c, ray.anchor)->center), add(V4(B(c, ray.anchor)->low), ray.offset));
}
static v3 to_ground(const Ctx *c, v3 v) {
    return V(dot(v, V4(c->s->ground_east)), dot(v, V4(c->s->ground_up)), dot(v, V4(c->s->ground_north)));
}
static v3 from_ground(const Ctx *c, v3 v) {
    return add(add(scl(V4(c->s->ground_east), v.x), scl(V4(c->s->ground_up), v.y)), scl(V4(c->s->ground_north), v.z));
}
static v3 ground_origin(const Ctx *c, Ray ray) {
    return add(to_ground(c, scl(camera_origin(c, ray), METRES_PER_AU)), V(0.0, c->s->ground_east.v[3], 0.0));
}
static v3 rotate_north(v3 p, R angle) {
    R cs = cos(angle), sn = sin(angle);
    return V(cs * p.x - sn * p.y, sn * p.x + cs * p.y, p.z);
}
static v3 rock_angles(R packed) {
    uint32_t bits = (uint32_t)packed;
    return scl(V((R)(bits & 255u), (R)((bits >> 8) & 255u), (R)((bits >> 16) & 255u)), TAU / 256.0);
}
static v3 rotate_east(v3 p, R angle) {
    R cs = cos(angle), sn = sin(angle);
    return V(p.x, cs * p.y - sn * p.z, sn * p.y + cs * p.z);
}
static v3 rotate_up(v3 p, R angle) {
    R cs = cos(angle), sn = sin(angle);
    return V(cs * p.x + sn * p.z, p.y, -sn * p.x + cs * p.z);
}
static v3 rock_to_local(v3 p, v3 a) { return rotate_east(rotate_up(rotate_north(p, -a.z), -a.y), -a.x); }
static v3 rock_from_local(v3 p, v3 a) { return rotate_north(rotate_up(rotate_east(p, a.x), a.y), a.z); }
static R comp(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* src/shaders/ground.wgsl:64-78 */
static void box_interval(v3 p, v3 d, v3 extent, R *out_near, R *out_far) {
    COUNT(EV_BOX);
    R nearv = -1.0e30, farv = 1.0e30;
    for (int axis = 0; axis < 3; ++axis) {
        R pa = comp(p, axis), da = comp(d, axis), ea = comp(extent, axis);
        if (fabs(da) < 1.0e-12) {
            if (fabs(pa) > ea) { *out_near = 1.0e30; *out_far = -1.0e30; return; }
        } else {
            R a = (-ea - pa) / da, b = (ea - pa) / da;
            nearv = mx(nearv, mn(a, b));
            farv = mn(farv, mx(a, b));
        }
    }
    *out_near = nearv;
    *out_far = farv;
}
/* src/shaders/ground.wgsl:82-121 */
static R rock_hit(v3 p, v3 d, v3 e, v3 *normal) {
    COUNT(EV_ROCK);
    v3 origin = dvv(p, e), direction = dvv(d, e);
    R nearv = -1.0e30, farv = 1.0e30;
    v3 near_normal = Vs(0.0), far_normal = Vs(0.0);
    for (uint32_t face = 0; face < 14u; ++face) {
        v3 n = Vs(0.0);
        R limit = 1.0;
        if (face < 6u) {
            R sg = (face & 1u) == 1u ? 1.0 : -1.0;
            if (face / 2u == 0u) n.x = sg; else if (face / 2u == 1u) n.y = sg; else n.z = sg;
        } else {
            uint32_t bits = face - 6u;
            n = V((bits & 1u) ? 1.0 : -1.0, (bits & 2u) ? 1.0 : -1.0, (bits & 4u) ? 1.0 : -1.0);
            limit = 1.65 + 0.08 * (R)(bits % 3u);
        }
        R denominator = dot(n, direction);
        R gap = limit - dot(n, origin);
        if (fabs(denominator) < 1.0e-12) {
            if (gap < 0.0) return 1.0e30;
        } else {
            R distance = gap / denominator;
            if (denominator < 0.0 && distance > nearv) { nearv = distance; near_normal = dvv(n, e); }
            if (denominator > 0.0 && distance < farv) { farv = distance; far_normal = dvv(n, e); }
        }
        if (nearv > farv) return 1.0e30;
    }
    if (farv <= 0.001) return 1.0e30;
    if (nearv > 0.001) { *normal = nrm(near_normal); return nearv; }
    *normal = nrm(far_normal);
    return farv;
}
/* src/shaders/terrain.wgsl:65-81 */
static v3 terrain_triangle(v3 o, v3 d, v3 a, v3 b, v3 cc) {
    COUNT(EV_TRIANGLE);
    v3 ab = sub(b, a), ac = sub(cc, a);
    v3 p = cross(d, ac);
    R determinant = dot(ab, p);
    if (fabs(determinant) < 1.0e-12) return V(0.0, 0.0, 1.0e30);
    R inverse = 1.0 / determinant;
    v3 relative = sub(o, a);
    R u = dot(relative, p) * inverse;
    v3 q = cross(relative, ab);
    R v = dot(d, q) * inverse;
    R distance = dot(ac, q) * inverse;
    if (u < -1.0e-6 || v < -1.0e-6 || u + v > 1.000001 || distance <= 0.001) return V(0.0, 0.0, 1.0e30);
    return V(u, v, distance);
}
/* src/shaders/forest.wgsl:56-136 */
static R trunk_hit(v3 p, v3 d, v3 e, v3 *normal) {
    v3 o = dvv(p, e), v = dvv(d, e);
    R best = 1.0e30;
    R a = v.x * v.x + v.z * v.z, b = o.x * v.x + o.z * v.z, cc = o.x * o.x + o.z * o.z - 1.0;
    R disc = b * b - a * cc;
    if (a > 1.0e-15 && disc >= 0.0) {
        for (uint32_t side = 0; side < 2u; ++side) {
            R t = (-b + sel(-sqrt(disc), sqrt(disc), side == 1u)) / a;
            v3 q = add(o, scl(v, t));
            if (t > 0.001 && t < best && fabs(q.y) <= 1.0) { best = t; *normal = nrm(dvv(V(q.x, 0.0, q.z), e)); }
        }
    }
    if (fabs(v.y) > 1.0e-15) {
        for (uint32_t side = 0; side < 2u; ++side) {
            R sy = sel(-1.0, 1.0, side == 1u);
            R t = (sy - o.y) / v.y;
            v3 q = add(o, scl(v, t));
            if (t > 0.001 && t < best && q.x * q.x + q.z * q.z <= 1.0) { best = t; *normal = V(0.0, sy, 0.0); }
        }
    }
    return best;
}
static R conifer_hit(v3 p, v3 d, v3 e, v3 *normal) {
    COUNT(EV_VEGETATION);
    R best = trunk_hit(add(p, V(0.0, e.y * 0.08, 0.0)), d, V(e.x * 0.065, e.y * 0.92, e.z * 0.065), normal);
    v3 o = dvv(p, e), v = dvv(d, e);
    for (uint32_t tier = 0; tier < 6u; ++tier) {
        R t = (R)tier;
        R base = -0.6 + t * 0.27;
        R height = mn(0.66 - t * 0.045, 1.0 - base);
        R radius = 0.94 - t * 0.148;
        v3 q = sub(o, V(sin(t * 4.1) * 0.022, base + height, cos(t * 3.7) * 0.022));
        R k = radius / height;
        R a = v.x * v.x + v.z * v.z - k * k * v.y * v.y;
        R b = q.x * v.x + q.z * v.z - k * k * q.y * v.y;
        R cc = q.x * q.x + q.z * q.z - k * k * q.y * q.y;
        R disc = b * b - a * cc;
        if (fabs(a) < 1.0e-12 || disc < 0.0) continue;
        for (uint32_t side = 0; side < 2u; ++side) {
            R distance = (-b + sel(-sqrt(disc), sqrt(disc), side == 1u)) / a;
            v3 h = add(q, scl(v, distance));
            if (distance <= 0.001 || distance >= best || h.y > 0.0 || h.y < -height) continue;
            R angle = atan2(h.z, h.x);
            R skirt = -h.y / height;
            R branches = 0.5 + 0.5 * sin(angle * (9.0 + t) + t * 3.7);
            R needle = vnoise(add(mul(add(p, scl(d, distance)), V(9.0, 17.0, 9.0)), Vs(t * 7.0)));
            if (skirt > 0.76 + 0.19 * branches || (skirt > 0.4 && needle < 0.26)) continue;
            best = distance;
            *normal = nrm(dvv(V(h.x, -k * k * h.y, h.z), e));
        }
    }
    return best;
}
static R grass_hit(v3 p, v3 d, v3 e, v3 *normal) {
    R best = 1.0e30;
    for (uint32_t blade = 0; blade < 9u; ++blade) {
        R i = (R)blade;
        R angle = i * 2.39996;
        v3 side = V(cos(angle), 0.0, sin(angle));
        v3 base = sub(scl(side, e.x * 0.15), V(0.0, e.y, 0.0));
        v3 tip = add(scl(side, e.x * (0.6 + 0.3 * sin(i * 3.0))), V(0.0, e.y * (0.5 + 0.5 * cos(i * 1.3)), 0.0));
        v3 width = scl(V(-side.z, 0.0, side.x), 0.035);
        v3 h = terrain_triangle(p, d, sub(base, width), add(base, width), tip);
        if (h.z < best) {
            v3 n = nrm(cross(width, sub(tip, base)));
            if (dot(n, d) > 0.0) n = neg(n);
            best = h.z;
            *normal = n;
        }
    }
    return best;
}

/* --- terrain (src/shaders/terrain.wgsl) --- */
typedef struct {
    uint32_t cells, width, compact, stride;
    uint64_t vertices, nodes, bounds, cache, vertex_cache;
} TL;
static TL terrain_layout(const Ctx *c) {
    TL t;
    uint32_t log2 = c->s->ground_counts[2] >> 8;
    t.cells = log2 != 0 ? (1u << log2) : 512u;
    t.width = t.cells + 1u;
    t.compact = t.cells > 1024u;
    t.stride = t.compact ? 2u : 6u;
    t.vertices = (uint64_t)t.width * t.width;
    uint64_t leaves = t.cells / 2u;
    t.nodes = (4u * leaves * leaves - 1u) / 3u;
    t.bounds = (t.compact ? 1u : 3u) * t.vertices;
    t.cache = t.bounds + 2u * t.nodes;
    t.vertex_cache = t.cache + 6u * t.nodes;
    return t;
}
static R terrain_scale(const Ctx *c) { return clampR((R)c->s->ground_up.v[3] / 200000.0, 0.0001, 1.0); }
static R terrain_coordinate(const Ctx *c, const TL *t, uint32_t i) {
    R u = 2.0 * (R)i / (R)t->cells - 1.0;
    return u * fabs(u) * 16000.0 * terrain_scale(c);
}
static R terrain_curvature(const Ctx *c, R px, R pz) {
    R r = c->s->ground_up.v[3];
    R d2 = px * px + pz * pz;
    return -d2 / (r + sqrt(mx(r * r - d2, 0.0)));
}
static v3 terrain_vertex(const Ctx *c, const TL *t, uint32_t x, uint32_t z) {
    uint64_t i = t->vertex_cache + ((uint64_t)z * t->width + x) * t->stride;
    if (t->compact) return V(terrain_coordinate(c, t, x), c->terrain[i], terrain_coordinate(c, t, z));
    return V(c->terrain[i], c->terrain[i + 1], c->terrain[i + 2]);
}
static v3 terrain_normal(const Ctx *c, const TL *t, uint32_t x, uint32_t z) {
    uint64_t i = t->vertex_cache + ((uint64_t)z * t->width + x) * t->stride;
    if (t->compact) {
        uint32_t packed = (uint32_t)c->terrain[i + 1];
        R nx = (R)(packed & 4095u) / 4095.0 * 2.0 - 1.0;
        R nz = (R)(packed >> 12u) / 4095.0 * 2.0 - 1.0;
        return nrm(V(nx, sqrt(mx(1.0 - (nx * nx + nz * nz), 1.0e-6)), nz));
    }
    return V(c->terrain[i + 3], c->terrain[i + 4], c->terrain[i + 5]);
}
static void terrain_interval(const Ctx *c, const TL *t, v3 o, v3 inv, uint32_t node, R *en, R *ex) {
    COUNT(EV_TERRAIN_NODE);
    const float *q = c->terrain + t->cache + (uint64_t)node * 6u;
    v3 lo = V(q[0], q[1], q[2]), hi = V(q[3], q[4], q[5]);
    int px = fabs(inv.x) >= 1.0e29, py = fabs(inv.y) >= 1.0e29, pz = fabs(inv.z) >= 1.0e29;
    if ((px && (o.x < lo.x || o.x > hi.x)) || (py && (o.y < lo.y || o.y > hi.y)) || (pz && (o.z < lo.z || o.z > hi.z))) {
        *en = 1.0e30;
        *ex = -1.0e30;
        return;
    }
    v3 a = mul(sub(lo, o), inv), b = mul(sub(hi, o), inv);
    R e0 = px ? -1.0e30 : mn(a.x, b.x), e1 = py ? -1.0e30 : mn(a.y, b.y), e2 = pz ? -1.0e30 : mn(a.z, b.z);
    R x0 = px ? 1.0e30 : mx(a.x, b.x), x1 = py ? 1.0e30 : mx(a.y, b.y), x2 = pz ? 1.0e30 : mx(a.z, b.z);
    *en = mx(e0, mx(e1, e2));
    *ex = mn(x0, mn(x1, x2));
}
static uint32_t compact_bits(uint32_t value) {
    uint32_t x = value & 0x55555555u;
    x = (x | (x >> 1u)) & 0x33333333u;
    x = (x | (x >> 2u)) & 0x0f0f0f0fu;
    x = (x | (x >> 4u)) & 0x00ff00ffu;
    return (x | (x >> 8u)) & 0x0000ffffu;
}
static uint32_t first_leading_bit(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

/* src/shaders/terrain.wgsl:152-234; returns distance, writes the normal */
static R terrain_query(const Ctx *c, v3 o, v3 d, R initial_limit, int occlusion, v3 *normal) {
    TL t = terrain_layout(c);
    R best = 1.0e30;
    R limit = initial_limit;
    v3 inv = V(sel(-1.0, 1.0, d.x >= 0.0) / mx(fabs(d.x), 1.0e-30), sel(-1.0, 1.0, d.y >= 0.0) / mx(fabs(d.y), 1.0e-30),
               sel(-1.0, 1.0, d.z >= 0.0) / mx(fabs(d.z), 1.0e-30));
    R rn, rx;
    terrain_interval(c, &t, o, inv, 0u, &rn, &rx);
    if (rn > rx || rx <= 0.001 || rn >= limit) return best;
    uint32_t sn[ORC_STACK];
    R st[ORC_STACK];
    sn[0] = 0u;
    st[0] = mx(rn, 0.0);
    uint32_t count = 1;
    while (count != 0u) {
        count -= 1u;
        uint32_t pending = sn[count];
        if (st[count] >= limit) continue;
        uint32_t level = first_leading_bit(3u * pending + 1u) / 2u;
        uint32_t size = t.cells >> level;
        uint32_t morton = pending - ((1u << (2u * level)) - 1u) / 3u;
        uint32_t nx = compact_bits(morton) * size, nz = compact_bits(morton >> 1u) * size;
        if (size == 2u) {
            for (uint32_t z = nz; z < nz + 2u; ++z) {
                for (uint32_t x = nx; x < nx + 2u; ++x) {
                    v3 a = terrain_vertex(c, &t, x, z), b = terrain_vertex(c, &t, x + 1u, z);
                    v3 cc = terrain_vertex(c, &t, x, z + 1u), e = terrain_vertex(c, &t, x + 1u, z + 1u);
                    v3 first = terrain_triangle(o, d, a, cc, b);
                    v3 second = terrain_triangle(o, d, e, b, cc);
                    if (first.z < limit) {
                        if (occlusion) return first.z;
                        *normal = nrm(add(add(scl(terrain_normal(c, &t, x, z), 1.0 - first.x - first.y),
                                              scl(terrain_normal(c, &t, x, z + 1u), first.x)),
                                          scl(terrain_normal(c, &t, x + 1u, z), first.y)));
                        best = first.z;
                        limit = first.z;
                    }
                    if (second.z < limit) {
                        if (occlusion) return second.z;
                        *normal = nrm(add(add(scl(terrain_normal(c, &t, x + 1u, z + 1u), 1.0 - second.x - second.y),
                                              scl(terrain_normal(c, &t, x + 1u, z), second.x)),
                                          scl(terrain_normal(c, &t, x, z + 1u), second.y)));
                        best = second.z;
                        limit = second.z;
                    }
                }
            }
        } else {
            uint32_t children[4];
            R distances[4];
            for (uint32_t child = 0; child < 4u; ++child) {
                children[child] = pending * 4u + child + 1u;
                R bn, bx;
                terrain_interval(c, &t, o, inv, children[child], &bn, &bx);
                distances[child] = (bn <= bx && bx > 0.001 && bn < limit) ? mx(bn, 0.0) : 1.0e30;
            }
            for (uint32_t i = 1; i < 4u; ++i) {
                uint32_t j = i;
                for (;;) {
                    if (j == 0u || distances[j - 1u] >= distances[j]) break;
                    R dt = distances[j - 1u];
                    uint32_t ct = children[j - 1u];
                    distances[j - 1u] = distances[j];
                    children[j - 1u] = children[j];
                    distances[j] = dt;
                    children[j] = ct;
                    j -= 1u;
                }
            }
            for (uint32_t i = 0; i < 4u; ++i) {
                if (distances[i] < limit && count < ORC_STACK) {
                    sn[count] = children[i];
                    st[count] = distances[i];
                    count += 1u;
                }
            }
        }
    }
    return best;
}

/* src/shaders/ground.wgsl:244-251 */
static R landscape_top(const Ctx *c) {
    uint32_t bits = c->s->ground_counts[3];
    float top;
    memcpy(&top, &bits, 4);
    R t = top;
    if (c->s->ground_counts[1] != 0u) {
        TL tl = terrain_layout(c);
        t = mx(t, c->terrain[tl.cache + 4u]);
    }
    return t;
}

/* src/shaders/ground.wgsl:123-242 */
static Hit landscape_query(const Ctx *c, Ray ray, R limit, int occlusion) {
    Hit closest = {limit, MISS, {0, 0, 0}};
    const Settings *s = c->s;
    if (s->ground_north.v[3] == 0.0f) return closest;
    v3 o = ground_origin(c, ray);
    v3 d = to_ground(c, ray.direction);
    R radius = s->ground_up.v[3];
    if (dot(o, o) > 4.0 * radius * radius) return sphere_hit(c, ray, (uint32_t)s->ground_north.v[3] - 1u);
    if (d.y >= 0.0 && o.y > landscape_top(c)) return closest;
    R a = dot(d, d);
    R b = -radius * d.y - dot(o, d);
    R cc = o.x * o.x + o.z * o.z + o.y * (2.0 * radius + o.y);
    R h = b * b - a * cc;
    if (h >= 0.0) {
        R root = sqrt(h);
        R q = b + sel(-root, root, b >= 0.0);
        if (q != 0.0) {
            R t0 = mn(cc / q, q / a), t1 = mx(cc / q, q / a);
            R distance = sel(t1, t0, t0 > 0.001);
            if (distance > 0.001 && distance / METRES_PER_AU < closest.distance) {
                if (occlusion) { Hit r = {distance / METRES_PER_AU, LOCAL_BASE, {0, 0, 0}}; return r; }
                v3 p = add(o, scl(d, distance));
                v3 n = nrm(V(p.x / radius, 1.0 + p.y / radius, p.z / radius));
                Hit r = {distance / METRES_PER_AU, LOCAL_BASE, from_ground(c, n)};
                closest = r;
            }
        }
    }
    if (s->ground_counts[1] != 0u) {
        R tl = mn(closest.distance, 1.0e18) * METRES_PER_AU;
        v3 tn = Vs(0.0);
        R tw = terrain_query(c, o, d, tl, occlusion, &tn);
        if (tw < tl && tw < 1.0e29) {
            Hit r = {tw / METRES_PER_AU, TERRAIN_ID, from_ground(c, tn)};
            closest = r;
            if (occlusion) return closest;
        }
    }
    for (uint32_t i = 0; i < s->ground_counts[0]; ++i) {
        const Prim *object = &c->objects[i];
        R kind = object->center.v[3];
        v3 oc = V4(object->center), e = V4(object->extent);
        if (kind < 0.0) {
            R bn, bx;
            box_interval(sub(o, oc), d, e, &bn, &bx);
            if (bn > bx || bx <= 0.001 || bn / METRES_PER_AU > closest.distance) i += (uint32_t)(-kind);
            continue;
        }
        v3 p = sub(o, oc), v = d, angles = Vs(0.0);
        if (kind == 2.25) {
            angles = rock_angles(object->extent.v[3]);
            p = rock_to_local(p, angles);
            v = rock_to_local(v, angles);
        } else {
            p = rotate_north(p, -(R)object->extent.v[3]);
            v = rotate_north(v, -(R)object->extent.v[3]);
        }
        R distance = 1.0e30;
        v3 n = Vs(0.0);
        if (kind > 2.5) {
            R bn, bx;
            box_interval(p, v, e, &bn, &bx);
            if (bn > bx || bx <= 0.001 || bn / METRES_PER_AU > closest.distance) continue;
            if (kind < 3.5) distance = conifer_hit(p, v, e, &n);
            else if (kind < 4.5) distance = grass_hit(p, v, e, &n);
            else distance = trunk_hit(p, v, e, &n);
        } else if (kind > 1.5) {
            R bn, bx;
            box_interval(p, v, e, &bn, &bx);
            if (bn > bx || bx <= 0.001 || bn / METRES_PER_AU > closest.distance) continue;
            distance = rock_hit(p, v, e, &n);
        } else if (kind > 0.5) {
            v3 sp = dvv(p, e), sd = dvv(v, e);
            R aa = dot(sd, sd), bb = dot(sp, sd), c2 = dot(sp, sp) - 1.0;
            R hh = bb * bb - aa * c2;
            if (hh >= 0.0) {
                R root = sqrt(hh);
                distance = (-bb - root) / aa;
                if (distance <= 0.001) distance = (-bb + root) / aa;
                n = nrm(dvv(add(p, scl(v, distance)), mul(e, e)));
            }
        } else {
            R bn, bx;
            box_interval(p, v, e, &bn, &bx);
            if (bn <= bx && bx > 0.001) {
                distance = sel(bx, bn, bn > 0.001);
                v3 point = dvv(add(p, scl(v, distance)), e);
                v3 ap = V(fabs(point.x), fabs(point.y), fabs(point.z));
                R sx = point.x > 0 ? 1.0 : (point.x < 0 ? -1.0 : 0.0);
                R sy = point.y > 0 ? 1.0 : (point.y < 0 ? -1.0 : 0.0);
                R sz = point.z > 0 ? 1.0 : (point.z < 0 ? -1.0 : 0.0);
                if (ap.x >= ap.y && ap.x >= ap.z) n.x = sx;
                else if (ap.y >= ap.z) n.y = sy;
                else n.z = sz;
            }
        }
        if (distance > 0.001 && distance < 1.0e29 && distance / METRES_PER_AU < closest.distance) {
            if (occlusion) { Hit r = {distance / METRES_PER_AU, LOCAL_BASE + 1u + i, {0, 0, 0}}; return r; }
            if (kind == 2.25) n = rock_from_local(n, angles);
            else n = rotate_north(n, object->extent.v[3]);
            Hit r = {distance / METRES_PER_AU, LOCAL_BASE + 1u + i, from_ground(c, n)};
            closest = r;
        }
    }
    return closest;
}
static Hit landscape_hit(const Ctx *c, Ray ray) { return landscape_query(c, ray, 1.0e30, 0); }

/* src/shaders/pathtrace.wgsl:173-185 */
static Hit closest_hit(const Ctx *c, Ray ray) {
    Hit closest = hit_miss();
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        if (c->s->ground_north.v[3] == (float)(i + 1u)) continue;
        Hit hit = sphere_hit(c, ray, i);
        if (hit.distance < closest.distance) closest = hit;
    }
    Hit local = landscape_hit(c, ray);
    if (local.distance < closest.distance) closest = local;
    return closest;
}

/* --- rings (src/shaders/rings.wgsl) --- */
static RingHit ring_miss(R limit) { RingHit r = {limit, MISS, {0, 0, 0}}; return r; }
static RingHit ring_hit(const Ctx *c, Ray ray, uint32_t index) {
    const Body *body = B(c, index);
    if (body->ring_params.v[0] <= body->ring_plane.v[3] || body->ring_params.v[1] <= 0.0f) return ring_miss(1.0e30);
    COUNT(EV_RING);
    v3 n = V4(body->ring_plane);
    R denom = dot(ray.direction, n);
    if (fabs(denom) < 1.0e-7) return ring_miss(1.0e30);
    v3 oc = relative_center(c, ray, index);
    R a = dot(ray.direction, ray.direction);
    v3 perpendicular = dvs(cross(ray.direction, cross(ray.direction, oc)), a);
    R along = -dot(perpendicular, n) / denom;
    R distance = dot(oc, ray.direction) / a + along;
    if (distance <= 0.0) return ring_miss(1.0e30);
    v3 point = add(perpendicular, scl(ray.direction, along));
    R radius = len(point) / (R)body->center.v[3];
    if (radius < (R)body->ring_plane.v[3] || radius > (R)body->ring_params.v[0]) return ring_miss(1.0e30);
    RingHit r = {distance, index, point};
    return r;
}
static RingHit closest_ring(const Ctx *c, Ray ray, R limit) {
    RingHit closest = ring_miss(limit);
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        RingHit hit = ring_hit(c, ray, i);
        if (hit.index != MISS && hit.distance < closest.distance) closest = hit;
    }
    return closest;
}
/* src/shaders/rings.wgsl:50-81; returns rgb in *color and tau */
static R ring_optics(const Ctx *c, uint32_t index, v3 point, v3 *color) {
    COUNT(EV_RING_OPTICS);
    const Body *body = B(c, index);
    R radius = len(point) / (R)body->center.v[3];
    R x = (radius - (R)body->ring_plane.v[3]) / ((R)body->ring_params.v[0] - (R)body->ring_plane.v[3]);
    R tau = body->ring_params.v[1];
    v3 col = V4(body->ring_color);
    if (body->ring_params.v[2] > 0.5f) {
        R broad = 0.5 + 0.5 * sin(43.0 * x + 1.8 * sin(17.0 * x));
        R fine = 0.5 + 0.5 * sin(1900.0 * x + 2.0 * sin(311.0 * x));
        R micro = 0.5 + 0.5 * sin(6100.0 * x);
        R density = 0.12;
        v3 tint = V(0.70, 0.68, 0.65);
        if (x >= 0.2792 && x < 0.6910) {
            density = 1.25 + 1.1 * broad;
            tint = Vs(1.0);
        } else if (x >= 0.6910 && x < 0.7649) {
            density = 0.025;
            tint = Vs(0.75);
        } else if (x >= 0.7649) {
            density = 0.48 + 0.22 * broad;
            tint = V(0.92, 0.94, 0.98);
        }
        density *= 0.65 + 0.45 * fine + 0.12 * micro;
        if (fabs(x - 0.9487) < 0.00262 || fabs(x - 0.99565) < 0.00034) density = 0.0;
        tau *= density;
        col = scl(mul(col, tint), 0.94 + 0.06 * broad);
    }
    if (index == c->ring_profile_body) {
        /* 512 radial bins over [inner, outer], clamped at both ends. */
        R bin = x * 512.0;
        uint32_t b = bin > 0.0 ? (uint32_t)bin : 0u;
        tau = c->ring_profile[b < 511u ? b : 511u];
    }
    if (color) *color = col;
    return tau;
}
static R ring_transmission(const Ctx *c, Ray ray, RingHit hit) {
    R mu = mx(fabs(dot(ray.direction, V4(B(c, hit.index)->ring_plane))), 1.0e-7);
    return exp(-ring_optics(c, hit.index, hit.point, 0) / mu);
}
/* src/shaders/rings.wgsl:143-166 */
static R one_minus_exp(R x) { return x < 1.0e-3 ? x * (1.0 - 0.5 * x) : 1.0 - exp(-x); }
static v3 ring_bsdf(v3 color, R tau, R outgoing, R incoming) {
    R mo = mx(fabs(outgoing), 1.0e-7), mi = mx(fabs(incoming), 1.0e-7);
    R a = tau / mo, b = tau / mi;
    if (outgoing * incoming < 0.0) {
        R delta = fabs(a - b);
        R integral = delta < 1.0e-3 ? 1.0 - 0.5 * delta : one_minus_exp(delta) / mx(delta, 1.0e-20);
        R factor = tau / (mo * mi) * exp(-mn(a, b)) * integral;
        return scl(color, factor / (4.0 * PI));
    }
    R factor = one_minus_exp(a + b) / (mo + mi);
    return scl(color, factor / (4.0 * PI));
}
static Ray ring_origin(const Ctx *c, RingHit hit, v3 direction) {
    const Body *body = B(c, hit.index);
    v3 n = V4(body->ring_plane);
    R sg = sel(-1.0, 1.0, dot(direction, n) >= 0.0);
    v3 point = sub(hit.point, scl(n, dot(hit.point, n)));
    Ray r = {add(point, scl(n, sg * (R)body->center.v[3] * 2.0e-6)), direction, hit.index};
    return r;
}
/* src/shaders/rings.wgsl:90-141 */
static R light_visibility(const Ctx *c, Ray ray, uint32_t light) {
    COUNT(EV_VISIBILITY);
    Hit surface = sphere_hit(c, ray, light);
    if (surface.index == MISS) return 0.0;
    const float host = c->s->ground_north.v[3];
    if (host == (float)(light + 1u) && closest_hit(c, ray).index != light) return 0.0;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        if (i == light || host == (float)(i + 1u)) continue;
        Hit hit = sphere_hit(c, ray, i);
        if (hit.index != MISS && (hit.distance < surface.distance || (hit.distance == surface.distance && i < light)))
            return 0.0;
    }
    Hit local = landscape_query(c, ray, surface.distance, 1);
    if (local.index != MISS && local.distance < surface.distance) return 0.0;
    R transmission = 1.0;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        RingHit hit = ring_hit(c, ray, i);
        if (hit.index != MISS && hit.distance < surface.distance) transmission *= ring_transmission(c, ray, hit);
    }
    return transmission;
}
static R air_light_visibility(const Ctx *c, Ray ray, uint32_t light) {
    if (c->s->ground_north.v[3] == 0.0f || ground_origin(c, ray).y <= landscape_top(c)) return light_visibility(c, ray, light);
    Hit surface = sphere_hit(c, ray, light);
    if (surface.index == MISS) return 0.0;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        if (i == light) continue;
        Hit hit = sphere_hit(c, ray, i);
        if (hit.index != MISS && (hit.distance < surface.distance || (hit.distance == surface.distance && i < light)))
            return 0.0;
    }
    R transmission = 1.0;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        RingHit hit = ring_hit(c, ray, i);
        if (hit.index != MISS && hit.distance < surface.distance) transmission *= ring_transmission(c, ray, hit);
    }
    return transmission;
}

/* --- ground materials, themes 0-3 and 5 (src/shaders/ground.wgsl:256-666) --- */
static R footprint_of(const Ctx *c, R distance) {
    return distance * 2.0 * (R)c->s->cam_forward.v[3] / mx((R)c->s->viewport.v[1], 1.0);
}
static v3 forest_floor(const Ctx *c, v3 p, R distance, v3 normal) {
    R broad = vnoise(add(scl(p, 0.045), V(7.0, 0.0, 13.0)));
    R middle = vnoise(scl(p, 0.7));
    R grain = vnoise(scl(p, 18.0));
    R detail = 1.0 - smooth(0.006, 0.055, footprint_of(c, distance));
    v3 moss = mixV(V(0.028, 0.054, 0.014), V(0.105, 0.14, 0.039), middle);
    v3 earth = mixV(V(0.055, 0.029, 0.013), V(0.17, 0.115, 0.057), middle);
    v3 color = mixV(earth, moss, smooth(0.35, 0.63, broad));
    R needles = pow(mx(0.0, sin(p.x * 92.0 + p.z * 137.0 + middle * 7.0)), 18.0) * smooth(0.50, 0.68, grain);
    color = mixV(color, V(0.22, 0.145, 0.067), needles * detail * 0.5);
    color = scl(color, 1.0 + (grain - 0.5) * 0.38 * detail);
    R slope = 1.0 - mx(normal.y, 0.0);
    v3 stone = mixV(V(0.095, 0.105, 0.10), V(0.24, 0.23, 0.20), middle);
    color = mixV(color, stone, smooth(0.10, 0.43, slope + (broad - 0.5) * 0.17));
    R center = 2.0 + 0.30 * p.z + 2.0 * sin(p.z * 0.055);
    R edge = fabs(p.x - center) + (middle - 0.5) * 0.55;
    R path = (1.0 - smooth(0.8, 2.1, edge)) * (1.0 - smooth(45.0, 95.0, p.z)) * smooth(-10.0, 0.0, p.z);
    v3 gravel = mixV(V(0.12, 0.095, 0.062), V(0.26, 0.23, 0.17), grain * detail + 0.5 * (1.0 - detail));
    return mixV(color, gravel, path * 0.85);
}
static uint32_t ground_theme(const Ctx *c) { return c->s->ground_counts[2] & 255u; } /* terrain.wgsl:22 */
static v3 desert_floor(R footprint, v3 p, v3 normal) { /* ground.wgsl:281-304 */
    R broad = vnoise(add(scl(p, 0.007), V(19.0, 2.0, 4.0)));
    R gravel = vnoise(scl(p, 2.8));
    R detail = 1.0 - smooth(0.025, 0.2, footprint);
    v3 sand = mixV(V(0.31, 0.21, 0.125), V(0.53, 0.40, 0.26), broad);
    v3 rock = mixV(V(0.18, 0.10, 0.055), V(0.38, 0.245, 0.145), broad);
    v3 color = mixV(sand, rock, smooth(0.08, 0.4, 1.0 - mx(normal.y, 0.0)));
    color = scl(color, 1.0 + (gravel - 0.5) * 0.32 * detail);
    R center = 2.8 + 0.55 * sin(p.z / 38.0) + 1.4 * sin(p.z / 210.0);
    R edge = fabs(p.x - center);
    R aa = mx(0.015, footprint * 0.5);
    R fade = (1.0 - smooth(0.25, 0.9, footprint)) * (1.0 - smooth(1000.0, 2200.0, fabs(p.z))) *
             (1.0 - smooth(0.12, 0.35, 1.0 - mx(normal.y, 0.0)));
    R worn = 1.0 - smooth(0.8, 1.2, edge);
    R rut_edge = fabs(edge - 0.68) + (gravel - 0.5) * 0.1 * detail;
    R tracks = 1.0 - smooth(0.12 - aa, 0.23 + aa, rut_edge);
    v3 packed = scl(sand, 0.72 + 0.10 * gravel * detail);
    color = mixV(color, scl(sand, 0.93), worn * fade * 0.35);
    return mixV(color, packed, tracks * fade * 0.8);
}
static v3 crater_floor(R footprint, v3 p, v3 normal) { /* ground.wgsl:306-317 */
    R broad = vnoise(add(scl(p, 0.004), V(3.0, 0.0, 8.0)));
    R middle = vnoise(scl(p, 0.09));
    R grain = vnoise(scl(p, 3.1));
    R detail = 1.0 - smooth(0.025, 0.2, footprint);
    R tone = mixR(0.11, 0.21, broad) * (0.86 + 0.28 * middle);
    tone *= 1.0 + 0.35 * smooth(0.15, 0.5, 1.0 - mx(normal.y, 0.0));
    tone *= 1.0 + (grain - 0.5) * 0.3 * detail;
    return V(tone, tone * 0.97, tone * 0.93);
}
static R ground_curvature(R r, R px, R pz) { /* terrain.wgsl:32-36 */
    R d2 = px * px + pz * pz;
    return -d2 / (r + sqrt(mx(r * r - d2, 0.0)));
}
#define ICE_FLOW 4.4
#define DUNE_WIND 0.61
#define DUNE_PAN_LEVEL 2.0
static R lookout_hill(R r) { /* ground.wgsl:322-325 */
    R u = mn(sqrt(r * r + 30.0 * 30.0) / 700.0, 1.0);
    return 180.0 * (1.0 - u) * (1.0 - u);
}
static R dune_cracks(R px, R py) { /* ground.wgsl:490-507 */
    R gx = floor(px), gy = floor(py), fx = px - gx, fy = py - gy;
    R d1 = 8.0, d2 = 8.0;
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
            R ox = (R)i, oy = (R)j;
            v3 h = hash33(V(gx + ox, gy + oy, 3.0));
            R dx = ox + 0.15 + 0.7 * h.x - fx, dy = oy + 0.15 + 0.7 * h.y - fy;
            R d = sqrt(dx * dx + dy * dy);
            if (d < d1) {
                d2 = d1;
                d1 = d;
            } else if (d < d2) {
                d2 = d;
            }
        }
    return d2 - d1;
}
static void ice_crevasse(R qx, R qz, R footprint, R out[2]) { /* ground.wgsl:333-362 */
    out[0] = out[1] = 0.0;
    R r = sqrt(qx * qx + qz * qz);
    R tension = smooth(650.0, 900.0, r) * (1.0 - smooth(1800.0, 2600.0, r));
    R strain = mx(smooth(0.47, 0.58, fbm(V(qx / 2600.0 + 4.0, 2.0, qz / 2600.0), 2)), tension) * smooth(600.0, 750.0, r);
    if (strain <= 0.0) return;
    R flx = sin((R)ICE_FLOW), flz = cos((R)ICE_FLOW);
    R along = qx * flx + qz * flz + 280.0 * (vnoise(V(qx / 1500.0, qz / 1500.0, 1.0)) - 0.5);
    R across = qx * flz - qz * flx + 120.0 * (vnoise(V(qx / 900.0, qz / 900.0, 5.0)) - 0.5);
    R spacing = 60.0;
    R row = floor(along / spacing);
    uint32_t key = orc_hash(((uint32_t)(int32_t)row * 0x8da6b343u) ^ 0xc3e5u);
    R length_m = 120.0 + 300.0 * (R)(key & 255u) / 255.0;
    R s = across / length_m + (R)((key >> 8) & 255u) / 255.0;
    R segment = floor(s);
    R lens = pow(mx(sin(PI * (s - segment)), 0.0), 0.7);
    uint32_t open = orc_hash(key ^ (uint32_t)(int32_t)segment);
    if ((open & 255u) < 70u) return;
    R centre = 0.5 + 0.3 * ((R)((open >> 8) & 255u) / 255.0 - 0.5);
    R half_width = 0.5 * (2.0 + 14.0 * (R)((open >> 16) & 255u) / 255.0) * lens * strain;
    R d = fabs(along / spacing - row - centre) * spacing + 0.8 * (vnoise(V(qx * 0.15, qz * 0.15, 9.0)) - 0.5);
    R aa = mx(footprint, 1.0e-3);
    out[0] = clampR((half_width - d) / aa + 0.5, 0.0, 1.0) * mn(1.0, 2.0 * half_width / aa);
    out[1] = 1.0 - clampR(d / mx(half_width, 1.0e-3), 0.0, 1.0);
}
static v3 alpine_floor(R footprint_m, R scale, R radius, v3 p, v3 normal) { /* ground.wgsl:365-428 */
    v3 q = dvs(V(p.x, p.y - ground_curvature(radius, p.x, p.z), p.z), scale);
    R hill = lookout_hill(sqrt(q.x * q.x + q.z * q.z));
    q.y -= hill;
    R footprint = footprint_m / scale;
    R detail = 1.0 - smooth(0.025, 0.2, footprint);
    R slope = 1.0 - mx(normal.y, 0.0);
    R broad = vnoise(add(scl(q, 0.004), V(11.0, 4.0, 1.0)));
    R middle = vnoise(scl(q, 0.2));
    R grain = vnoise(scl(q, 4.0));
    R flx = sin((R)ICE_FLOW), flz = cos((R)ICE_FLOW);
    R along = q.x * flx + q.z * flz, across = q.x * flz - q.z * flx;
    R crust = smooth(0.45, 0.6, fbm(V(along * 0.004, 2.0, across * 0.012), 3));
    v3 snow = mixV(V(0.86, 0.88, 0.91), V(0.76, 0.8, 0.86), crust);
    snow = scl(snow, 0.97 + 0.05 * middle);
    R bare = smooth(0.47, 0.53, fbm(add(scl(q, 0.0016), V(3.0, 0.0, 8.0)), 4) + 0.1 * (broad - 0.5));
    v3 ice = mixV(V(0.36, 0.55, 0.7), V(0.52, 0.67, 0.78), middle);
    R fracture_width = 0.03 + footprint / 4.0;
    R fractures = (1.0 - smooth(0.0, fracture_width, dune_cracks(q.x / 4.0 + 9.0, q.z / 4.0 + 9.0))) * 0.03 / fracture_width;
    ice = mixV(ice, V(0.82, 0.86, 0.9), mn(fractures, 1.0) * 0.7);
    ice = scl(ice, 1.0 + (grain - 0.5) * 0.12 * detail);
    v3 glacier = mixV(snow, ice, bare);
    v3 debris = mixV(V(0.12, 0.115, 0.11), V(0.26, 0.24, 0.22), grain);
    R dirty = smooth(6.0, 30.0, q.y) * (1.0 - smooth(30.0, 70.0, q.y));
    glacier = mixV(glacier, mixV(debris, glacier, 0.45), dirty * smooth(0.4, 0.7, broad) * 0.8);
    R wall = smooth(0.35, 0.6, slope) * (1.0 - smooth(20.0, 45.0, q.y));
    R depth = smooth(4.0, -4.0, q.y);
    v3 blue = mixV(V(0.32, 0.55, 0.72), V(0.1, 0.25, 0.42), depth);
    glacier = mixV(glacier, scl(blue, 0.9 + 0.2 * middle), wall);
    R cr[2];
    ice_crevasse(q.x, q.z, footprint, cr);
    R crack_x = cr[0] * (1.0 - smooth(10.0, 25.0, q.y)), crack_y = cr[1] * (1.0 - smooth(10.0, 25.0, q.y));
    R lip = (1.0 - smooth(0.0, 0.12, crack_y)) * crack_x;
    v3 inside = mixV(V(0.3, 0.55, 0.74), V(0.02, 0.06, 0.12), smooth(0.1, 0.8, crack_y));
    glacier = mixV(glacier, inside, crack_x);
    glacier = mixV(glacier, V(0.9, 0.92, 0.95), lip * 0.5);
    R mountain = mx(smooth(25.0, 70.0, q.y + 30.0 * (broad - 0.5)), smooth(4.0, 15.0, hill));
    R joints = 1.0 - smooth(0.0, 0.06 + footprint / 3.0, dune_cracks((q.x + q.y * 0.5) / 3.0, q.z / 3.0));
    v3 granite = mixV(V(0.09, 0.085, 0.08), V(0.24, 0.215, 0.19), broad);
    granite = scl(granite, 0.85 + 0.3 * middle - 0.35 * joints * (1.0 - smooth(0.1, 1.0, footprint)));
    R lichen = smooth(0.68, 0.78, vnoise(scl(q, 0.9))) * (1.0 - smooth(0.6, 0.9, slope));
    granite = mixV(granite, V(0.3, 0.17, 0.06), lichen * 0.5 * detail);
    R cover = 1.0 - smooth(0.17, 0.33, slope + (broad - 0.5) * 0.18 + (middle - 0.5) * 0.08);
    R scoured = smooth(0.52, 0.62, fbm(add(scl(q, 0.012), V(2.0, 0.0, 6.0)), 4)) * smooth(25.0, 90.0, hill) *
                (1.0 - smooth(150.0, 172.0, hill));
    R cover_hill = cover * (1.0 - scoured);
    v3 peak = mixV(granite, snow, cover_hill);
    v3 color = mixV(glacier, peak, mountain);
    return scl(color, 1.0 + (grain - 0.5) * 0.08 * detail);
}
static void alpine_gradient(R footprint, v3 p, v3 n, R g[2]) { /* ground.wgsl:430-461 */
    R snowy = 1.0 - smooth(0.3, 0.5, 1.0 - n.y);
    R flx = sin((R)ICE_FLOW), flz = cos((R)ICE_FLOW), sx = flz, sz = -flx;
    R gx = 0.0, gz = 0.0, scale = 1.2, strength = 0.2;
    for (uint32_t octave = 0; octave < 3u; ++octave) {
        R fade = 1.0 - smooth(0.08 * scale, 0.25 * scale, footprint);
        if (fade > 0.0) {
            R a = (p.x * flx + p.z * flz) / (6.0 * scale), cc = (p.x * sx + p.z * sz) / scale;
            R e = 0.08, o = (R)octave * 13.0;
            R da = vnoise(V(a + e, o, cc)) - vnoise(V(a - e, o, cc));
            R dc = vnoise(V(a, o, cc + e)) - vnoise(V(a, o, cc - e));
            R crest = 1.0 + 1.5 * (1.0 - fabs(vnoise(V(a, o, cc)) * 2.0 - 1.0));
            R w = strength * fade * crest * scale;
            gx += (flx * da / (6.0 * scale) + sx * dc / scale) / (2.0 * e) * w;
            gz += (flz * da / (6.0 * scale) + sz * dc / scale) / (2.0 * e) * w;
        }
        scale *= 6.0;
        strength *= 0.6;
    }
    g[0] = gx * snowy;
    g[1] = gz * snowy;
}
static v3 dune_floor(R footprint_m, R scale, R radius, v3 p, v3 normal) { /* ground.wgsl:509-569 */
    v3 q = dvs(V(p.x, p.y - ground_curvature(radius, p.x, p.z), p.z), scale);
    R footprint = footprint_m / scale;
    R detail = 1.0 - smooth(0.02, 0.15, footprint);
    R slope = 1.0 - mx(normal.y, 0.0);
    R broad = vnoise(add(scl(q, 0.0015), V(1.0, 7.0, 3.0)));
    R middle = vnoise(scl(q, 0.05));
    R rise = smooth(6.0, 160.0, q.y + 40.0 * (broad - 0.5));
    v3 plinth = mixV(V(0.3, 0.095, 0.038), V(0.36, 0.12, 0.045), broad);
    v3 upper = mixV(V(0.42, 0.155, 0.06), V(0.47, 0.19, 0.072), broad);
    v3 sand = scl(mixV(plinth, upper, rise), 0.94 + 0.08 * middle);
    R flat = sqrt(normal.x * normal.x + normal.z * normal.z);
    if (flat > 1.0e-3) {
        R dx = -normal.x / flat, dz = -normal.z / flat;
        R across = q.x * -dz + q.z * dx, along = q.x * dx + q.z * dz;
        R tongues = fbm(V(across * 0.18, along * 0.012, 5.0), 3);
        R slip = smooth(0.32, 0.45, slope);
        sand = mixV(sand, scl(mul(sand, V(1.06, 1.12, 1.16)), 0.86 + 0.28 * tongues), slip);
    }
    R lag = smooth(0.6, 0.76, fbm(add(mul(q, V(0.013, 0.0, 0.004)), V(9.0, 0.0, 2.0)), 4)) *
            (1.0 - smooth(0.04, 0.16, slope)) * (1.0 - smooth(10.0, 45.0, q.y));
    sand = mixV(sand, V(0.13, 0.065, 0.05), lag * 0.55);
    R corridor = (1.0 - smooth(5.0, 16.0, q.y + 6.0 * (broad - 0.5))) * (1.0 - smooth(0.03, 0.1, slope));
    v3 gravel = mixV(V(0.39, 0.21, 0.115), V(0.45, 0.28, 0.16), vnoise(scl(q, 0.3)));
    R pebbles = smooth(0.72, 0.8, vnoise(scl(q, 7.0))) * detail;
    sand = mixV(sand, scl(gravel, 1.0 - 0.45 * pebbles), corridor * 0.6);
    static const float pans[2][3] = {{842.6f, 707.1f, 560.0f}, {4346.7f, 1164.7f, 900.0f}};
    R pan = 0.0;
    for (uint32_t i = 0; i < 2u; ++i) {
        R ex = q.x - (R)pans[i][0], ez = q.z - (R)pans[i][1];
        pan = mx(pan, 1.0 - smooth((R)pans[i][2] * 1.2, (R)pans[i][2] * 1.35 + 140.0, sqrt(ex * ex + ez * ez)));
    }
    pan *= 1.0 - smooth(DUNE_PAN_LEVEL + 0.25, DUNE_PAN_LEVEL + 1.5, q.y);
    if (pan > 0.0) {
        R tone = fbm(add(scl(q, 0.012), V(4.0, 0.0, 1.0)), 4);
        v3 clay = mixV(V(0.4, 0.35, 0.29), V(0.6, 0.56, 0.49), smooth(0.3, 0.7, tone));
        clay = scl(clay, 0.82 + 0.18 * smooth(0.35, 1.0, pan));
        R fine_width = 0.04 + footprint / 0.9;
        R cracks = (1.0 - smooth(0.0, fine_width, dune_cracks(q.x / 0.9, q.z / 0.9))) * 0.04 / fine_width;
        R wide_width = 0.03 + footprint / 6.0;
        R large = (1.0 - smooth(0.0, wide_width, dune_cracks(q.x / 6.0 + 17.0, q.z / 6.0 + 17.0))) * 0.03 / wide_width;
        clay = scl(clay, 1.0 - 0.6 * mn(cracks, 1.0) - 0.45 * mn(large, 1.0));
        clay = scl(clay, 0.92 + 0.16 * vnoise(scl(q, 2.3)) * (1.0 - smooth(0.04, 0.25, footprint)));
        sand = mixV(sand, clay, pan);
    }
    return scl(sand, 1.0 + (vnoise(scl(q, 9.0)) - 0.5) * 0.12 * detail);
}
static void dune_ripples(v3 p, R wavelength, R warp, R amplitude, R g[2]) { /* ground.wgsl:571-581 */
    R wx = sin((R)DUNE_WIND), wz = cos((R)DUNE_WIND), sx = wz, sz = -wx;
    R bend = warp * (vnoise(V(p.x / (9.0 * wavelength), p.z / (9.0 * wavelength), 1.0)) +
                     0.35 * vnoise(V(p.x / (2.5 * wavelength), p.z / (2.5 * wavelength), 7.0)));
    R phase = (p.x * wx + p.z * wz) / wavelength + bend + 0.4 * sin((p.x * sx + p.z * sz) / (3.0 * wavelength));
    R theta = TAU * phase;
    R slope = cos(theta + 0.65 * sin(theta)) * (1.0 + 0.65 * cos(theta));
    R patches = smooth(0.25, 0.6, vnoise(V(p.x / (14.0 * wavelength), p.z / (14.0 * wavelength), 3.0)));
    R k = slope * amplitude * TAU / wavelength * patches;
    g[0] = wx * k;
    g[1] = wz * k;
}
static void dune_gradient(R footprint, R scale, R radius, v3 p, v3 n, R g[2]) { /* ground.wgsl:583-595 */
    R gentle = 1.0 - smooth(0.22, 0.38, 1.0 - n.y);
    R qy = (p.y - ground_curvature(radius, p.x, p.z)) / scale;
    R dry = smooth(DUNE_PAN_LEVEL + 0.4, DUNE_PAN_LEVEL + 1.5, qy);
    R gx = 0.0, gz = 0.0, r[2];
    R fine = 1.0 - smooth(0.03, 0.07, footprint);
    if (fine > 0.0) {
        dune_ripples(p, 0.12, 2.5, 0.006, r);
        gx += r[0] * fine;
        gz += r[1] * fine;
    }
    R coarse = (1.0 - smooth(0.15, 0.4, footprint)) * (1.0 - smooth(9.0, 16.0, qy));
    if (coarse > 0.0) {
        dune_ripples(add(p, Vs(31.0)), 0.8, 2.0, 0.03, r);
        gx += r[0] * coarse;
        gz += r[1] * coarse;
    }
    g[0] = gx * gentle * dry;
    g[1] = gz * gentle * dry;
}
static v3 landscape_point(const Ctx *c, Ray ray, Hit hit) {
    return add(ground_origin(c, ray), scl(to_ground(c, ray.direction), hit.distance * METRES_PER_AU));
}
/* Ground material at surface-frame point p (metres) of landscape element
 * `index`, seen from `distance` metres, with surface-frame normal n
 * (ground.wgsl:597-640). */
static v3 ground_albedo(const Ctx *c, v3 p, R distance, v3 n, uint32_t index) {
    if (index == LOCAL_BASE || index == TERRAIN_ID) {
        uint32_t theme = ground_theme(c);
        if (theme == 1u) return desert_floor(footprint_of(c, distance), p, n);
        if (theme == 2u) return crater_floor(footprint_of(c, distance), p, n);
        if (theme == 3u) return alpine_floor(footprint_of(c, distance), terrain_scale(c), (R)c->s->ground_up.v[3], p, n);
        if (theme == 5u) return dune_floor(footprint_of(c, distance), terrain_scale(c), (R)c->s->ground_up.v[3], p, n);
        return forest_floor(c, p, distance, n);
    }
    const Prim *object = &c->objects[index - LOCAL_BASE - 1u];
    v3 tint = V4(object->albedo);
    R kind = object->albedo.v[3];
    if (kind == 0.0) return tint;
    R middle = vnoise(scl(p, 1.7));
    R fine = mixR(vnoise(scl(p, 12.0)), 0.5, smooth(18.0, 110.0, distance));
    if (kind > 4.5) return scl(tint, 0.65 + 0.6 * middle);
    if (kind > 2.5) {
        v3 q = dvv(rotate_north(sub(p, V4(object->center)), -(R)object->extent.v[3]), V4(object->extent));
        if (kind > 3.5 || (q.y < -0.35 && sqrt(q.x * q.x + q.z * q.z) < 0.09)) {
            R bark = vnoise(mul(p, V(25.0, 1.5, 25.0)));
            return scl(V(0.13, 0.073, 0.032), 0.5 + 0.85 * bark);
        }
        return scl(tint, 0.7 + 0.55 * middle + 0.2 * fine);
    }
    R strata = sin(p.y * 8.0 + 2.5 * middle);
    R lichen = smooth(0.52, 0.76, middle) * mx(n.y, 0.0);
    v3 stone = scl(tint, 0.7 + 0.45 * fine + 0.07 * strata);
    if (ground_theme(c) != 0u) return stone; /* only the forest's rocks carry moss and lichen */
    return mixV(stone, V(0.11, 0.145, 0.037), lichen * 0.65);
}
static v3 landscape_albedo(const Ctx *c, Ray ray, Hit hit) {
    return ground_albedo(c, landscape_point(c, ray, hit), hit.distance * METRES_PER_AU, to_ground(c, hit.normal),
                         hit.index);
}
/* Shading normal (world frame) at surface-frame point p of landscape element
 * `index`, seen from `distance` metres, with world-frame geometric normal
 * (ground.wgsl:641-648). */
static v3 ground_shading_normal(const Ctx *c, v3 p, R distance, v3 normal, uint32_t index) {
    if (index != LOCAL_BASE && index != TERRAIN_ID) return normal;
    uint32_t theme = ground_theme(c);
    if (theme == 3u || theme == 5u) { /* ground.wgsl:642-643 */
        v3 n = to_ground(c, normal);
        R g[2];
        if (theme == 3u) alpine_gradient(footprint_of(c, distance), p, n, g);
        else dune_gradient(footprint_of(c, distance), terrain_scale(c), (R)c->s->ground_up.v[3], p, n, g);
        return nrm(add(normal, from_ground(c, V(-g[0], 0.0, -g[1]))));
    }
    R strength = 0.11 * (1.0 - smooth(8.0, 65.0, distance));
    R dx = vnoise(scl(add(p, V(0.025, 0.0, 0.0)), 6.0)) - vnoise(scl(sub(p, V(0.025, 0.0, 0.0)), 6.0));
    R dz = vnoise(scl(add(p, V(0.0, 0.0, 0.025)), 6.0)) - vnoise(scl(sub(p, V(0.0, 0.0, 0.025)), 6.0));
    return nrm(add(normal, scl(from_ground(c, V(-dx, 0.0, -dz)), strength)));
}
static v3 landscape_shading_normal(const Ctx *c, Ray ray, Hit hit) {
    if (hit.index != LOCAL_BASE && hit.index != TERRAIN_ID) return hit.normal;
    return ground_shading_normal(c, landscape_point(c, ray, hit), hit.distance * METRES_PER_AU, hit.normal,
                                 hit.index);
}
static Ray landscape_outgoing(const Ctx *c, Ray ray, Hit hit) {
    v3 p = landscape_point(c, ray, hit);
    if (hit.index == LOCAL_BASE) {
        R radius = c->s->ground_up.v[3];
        R d2 = p.x * p.x + p.z * p.z;
        if (p.y > -radius && d2 < radius * radius) p.y = -d2 / (radius + sqrt(radius * radius - d2));
    }
    p = add(p, scl(to_ground(c, hit.normal), 0.005));
    p.y -= (R)c->s->ground_east.v[3];
    Ray r = {dvs(from_ground(c, p), METRES_PER_AU), hit.normal, MISS};
    return r;
}

/* --- atmosphere (src/shaders/pathtrace.wgsl:927-1234) --- */
static v3 atmo_c(const Ctx *c) { return V4(c->s->atmo_center); }
static R host_radius(const Ctx *c) { return c->s->atmo_center.v[3]; }
static R atmosphere_top(const Ctx *c) { return (R)c->s->atmo_center.v[3] + (R)c->s->atmo_params.v[2]; }
static R scale_height(const Ctx *c) { return c->s->atmo_params.v[1]; }
static R shell_exit(const Ctx *c, v3 origin, v3 dir) {
    v3 oc = sub(origin, atmo_c(c));
    R b = dot(oc, dir);
    R top = atmosphere_top(c);
    R cc = dot(oc, oc) - top * top;
    R disc = b * b - cc;
    if (disc < 0.0) return 0.0;
    return -b + sqrt(disc);
}
/* Whether Mie has its own scale height (extension; the WGSL has one H). */
static int two_heights(const Ctx *c) { return c->mie_height > 0.0; }
/* wgsl:958-971. Under two scale heights *mie receives the Mie column. */
static R density_integral(const Ctx *c, v3 origin, v3 dir, uint32_t steps, R *mie) {
    R distance = shell_exit(c, origin, dir);
    if (mie) *mie = 0.0;
    if (distance <= 0.0) return 0.0;
    if (counting) local_counters[stage][EV_DENSITY_STEP] += steps;
    R dt = distance / (R)steps;
    R inv_h = 1.0 / scale_height(c);
    R inv_hm = two_heights(c) ? 1.0 / c->mie_height : 0.0;
    R r = host_radius(c);
    R sum = 0.0, sum_m = 0.0;
    for (uint32_t i = 0; i < steps; ++i) {
        v3 p = add(origin, scl(dir, ((R)i + 0.5) * dt));
        R h = mx(len(sub(p, atmo_c(c))) - r, 0.0);
        sum += exp(-h * inv_h);
        if (two_heights(c)) sum_m += exp(-h * inv_hm);
    }
    if (mie && two_heights(c)) *mie = sum_m * dt;
    return sum * dt;
}
static v3 host_offset(const Ctx *c, Ray origin) {
    if (origin.anchor == MISS) return sub(origin.offset, atmo_c(c));
    const Body *body = B(c, origin.anchor);
    return add(add(sub(V4(body->center), atmo_c(c)), V4(body->low)), origin.offset);
}
static int inside_air(const Ctx *c, Ray origin) {
    if (c->s->atmo_params.v[3] < 0.5f) return 0;
    v3 oc = host_offset(c, origin);
    R top = atmosphere_top(c);
    return dot(oc, oc) < top * top;
}
static v3 beta_r(const Ctx *c) { return V4(c->s->atmo_rayleigh); }
static R beta_m(const Ctx *c) { return c->s->atmo_rayleigh.v[3]; }
/* Optical depth of a Rayleigh and a Mie column; with one scale height the
 * columns are equal and the depth is (beta_r + beta_m) * column (wgsl:996). */
static v3 air_depth(const Ctx *c, R column, R mie_column) {
    if (two_heights(c)) return add(scl(beta_r(c), column), Vs(beta_m(c) * mie_column));
    return scl(add(beta_r(c), Vs(beta_m(c))), column);
}
static v3 air_transmittance(const Ctx *c, Ray origin, v3 dir) {
    if (!inside_air(c, origin)) return Vs(1.0);
    v3 p = add(atmo_c(c), host_offset(c, origin));
    R mie;
    R column = density_integral(c, p, dir, c->sun_steps, &mie);
    return expV(neg(air_depth(c, column, mie)));
}
static R rayleigh_phase(R ct) { return 3.0 / (16.0 * PI) * (1.0 + ct * ct); }
static R mie_phase(R ct, R g) {
    R g2 = g * g;
    return (1.0 - g2) / (4.0 * PI * pow(1.0 + g2 - 2.0 * g * ct, 1.5));
}
static Ray camera_ray(v3 o, v3 d) { Ray r = {o, d, MISS}; return r; }
static uint32_t sample_reflector(const Ctx *c, v3 origin, uint32_t *rng, R *inverse_probability) {
    Ray ray = camera_ray(origin, V(0.0, 0.0, -1.0));
    uint32_t host = (uint32_t)c->s->atmo_params.v[3] - 1u;
    uint32_t selected = MISS;
    R selected_weight = 0.0, total = 0.0;
    for (uint32_t b = 0; b < c->s->counts[0]; ++b) {
        const Body *body = B(c, b);
        if (b == host || body->material.v[3] > 0.5f) continue;
        R weight = cone_width(c, ray, b) * maxc(V4(body->material));
        if (weight <= 0.0) continue;
        total += weight;
        if (random_(rng) * total < weight) {
            selected = b;
            selected_weight = weight;
        }
    }
    *inverse_probability = total / mx(selected_weight, 1.0e-30);
    return selected;
}
static v3 reflected_radiance(const Ctx *c, Ray ray, uint32_t index, uint32_t *rng) {
    Hit hit = sphere_hit(c, ray, index);
    if (hit.index == MISS) return Vs(0.0);
    const Body *body = B(c, index);
    v3 albedo = surface_albedo(body, hit.normal);
    Ray outgoing = {scl(hit.normal, (R)body->center.v[3] * (1.0 + 2.0e-6)), hit.normal, index};
    v3 radiance = Vs(0.0);
    for (uint32_t slot = 0; slot < light_count(c); ++slot) {
        uint32_t s = light_index(c, slot);
        if (B(c, s)->material.v[3] < 0.5f) continue;
        LightSample light = sample_light(c, outgoing, s, rng);
        R cosine = mx(dot(hit.normal, light.direction), 0.0);
        if (cosine > 0.0) {
            outgoing.direction = light.direction;
            radiance = add(radiance, scl(mul(albedo, V4(B(c, s)->material)),
                                         cosine * light_visibility(c, outgoing, s) / (PI * light.pdf)));
        }
    }
    return radiance;
}
static R brightest_risen_light(const Ctx *c, v3 origin) {
    Ray ray = camera_ray(origin, V(0.0, 0.0, -1.0));
    R brightest = 0.0;
    for (uint32_t slot = 0; slot < light_count(c); ++slot) {
        uint32_t b = light_index(c, slot);
        if (B(c, b)->material.v[3] < 0.5f) continue;
        if (dot(nrm(relative_center(c, ray, b)), V4(c->s->ground_up)) <= 0.0) continue;
        brightest = mx(brightest, maxc(V4(B(c, b)->material)) * TAU * cone_width(c, ray, b));
    }
    return brightest;
}
/* src/shaders/pathtrace.wgsl:1092-1213; returns the in-scatter */
static v3 atmosphere(const Ctx *c, v3 origin, v3 dir, R limit, uint32_t *rng) {
    v3 inscatter = Vs(0.0);
    if (c->s->atmo_params.v[3] < 0.5f) return inscatter;
    R exit = shell_exit(c, origin, dir);
    R t_max = mn(limit, exit);
    if (t_max <= 0.0) return inscatter;
    v3 br = beta_r(c);
    R bm = beta_m(c);
    v3 ext = add(br, Vs(bm));
    R g = c->s->atmo_params.v[0];
    R r = host_radius(c);
    R H = scale_height(c);
    uint32_t steps = (uint32_t)ceil(t_max / (0.25 * H));
    if (steps < 2u) steps = 2u;
    if (steps > c->view_steps) steps = c->view_steps;
    R dt = t_max / (R)steps;
    uint32_t jitter_rng = orc_hash(*rng ^ 0x3c6ef372u);
    R jd = random_(&jitter_rng);
    R jitter = (steps < c->view_steps || limit < exit) ? jd : 0.5;
    uint32_t reflected_rng = orc_hash(*rng ^ 0xa511e9b3u);
    R inverse_probability;
    uint32_t reflector = sample_reflector(c, origin, &reflected_rng, &inverse_probability);
    v3 reflected = Vs(0.0), reflected_direction = dir;
    R reflected_rayleigh = 0.0, reflected_mie = 0.0;
    if (reflector != MISS) {
        Ray reflected_ray = camera_ray(origin, dir);
        LightSample light = sample_light(c, reflected_ray, reflector, &reflected_rng);
        reflected_ray.direction = light.direction;
        reflected = scl(reflected_radiance(c, reflected_ray, reflector, &reflected_rng), inverse_probability / light.pdf);
        reflected_direction = light.direction;
        R ct = dot(dir, light.direction);
        reflected_rayleigh = rayleigh_phase(ct);
        reflected_mie = mie_phase(ct, g);
    }
    if (reflected.x > 0.0 || reflected.y > 0.0 || reflected.z > 0.0) {
        R sun = brightest_risen_light(c, origin);
        if (sun > 0.0) {
            R ratio = maxc(reflected) / sun;
            R minimum = planetshine_thinning ? 1.0 / sqrt(clampR((R)c->accumulated + 1.0, 1.0, 4096.0)) : 1.0;
            R keep = clampR(ratio * 1.0e3, minimum, 1.0);
            reflected = random_(&reflected_rng) < keep ? dvs(reflected, keep) : Vs(0.0);
        }
    }
    int planetshine = reflected.x > 0.0 || reflected.y > 0.0 || reflected.z > 0.0;
    v3 t_view = Vs(1.0);
    for (uint32_t i = 0; i < steps; ++i) {
        COUNT(EV_ATMO_CELL);
        v3 p = add(origin, scl(dir, ((R)i + jitter) * dt));
        R altitude = mx(len(sub(p, atmo_c(c))) - r, 0.0);
        R density = exp(-altitude / H);
        R column = density * dt;
        v3 tau, step_t, segment;
        /* Scattering per step = (beta_r scale_r, beta_m scale_m) x segment:
         * one scale height keeps wgsl:1163-1171 with unit scales; two give
         * each species its own density, the segment then being the
         * transmittance integral over the step's length. */
        R scale_r = 1.0, scale_m = 1.0;
        if (two_heights(c)) {
            R density_m = exp(-altitude / c->mie_height);
            tau = add(scl(br, column), Vs(bm * (density_m * dt)));
            step_t = expV(neg(tau));
            segment.x = tau.x < 1.0e-3 ? dt * (1.0 - 0.5 * tau.x) : (1.0 - step_t.x) * dt / tau.x;
            segment.y = tau.y < 1.0e-3 ? dt * (1.0 - 0.5 * tau.y) : (1.0 - step_t.y) * dt / tau.y;
            segment.z = tau.z < 1.0e-3 ? dt * (1.0 - 0.5 * tau.z) : (1.0 - step_t.z) * dt / tau.z;
            scale_r = density;
            scale_m = density_m;
        } else {
            tau = scl(ext, column);
            step_t = expV(neg(tau));
            segment.x = tau.x < 1.0e-3 ? column * (1.0 - 0.5 * tau.x) : (1.0 - step_t.x) / mx(ext.x, 1.0e-20);
            segment.y = tau.y < 1.0e-3 ? column * (1.0 - 0.5 * tau.y) : (1.0 - step_t.y) / mx(ext.y, 1.0e-20);
            segment.z = tau.z < 1.0e-3 ? column * (1.0 - 0.5 * tau.z) : (1.0 - step_t.z) / mx(ext.z, 1.0e-20);
        }
        v3 br_step = scl(br, scale_r);
        R bm_step = bm * scale_m;
        for (uint32_t slot = 0; slot < light_count(c); ++slot) {
            uint32_t b = light_index(c, slot);
            if (B(c, b)->material.v[3] < 0.5f) continue;
            Ray light_ray = camera_ray(p, dir);
            uint32_t ls_count = eclipse_light_samples(c, light_ray, b);
            for (uint32_t ls = 0; ls < ls_count; ++ls) {
                LightSample light = sample_light_stratum(c, light_ray, b, ls, ls_count, rng);
                light_ray.direction = light.direction;
                R visibility = air_light_visibility(c, light_ray, b);
                if (visibility == 0.0) continue;
                v3 e_sun = scl(V4(B(c, b)->material), visibility / ((R)ls_count * light.pdf));
                R sun_mie;
                R sun_column = density_integral(c, p, light.direction, c->sun_steps, &sun_mie);
                v3 sun_t = expV(neg(air_depth(c, sun_column, sun_mie)));
                R ct = dot(dir, light.direction);
                v3 beta_phase = add(scl(br_step, rayleigh_phase(ct)), Vs(bm_step * mie_phase(ct, g)));
                inscatter = add(inscatter, mul(mul(mul(mul(t_view, segment), beta_phase), e_sun), sun_t));
            }
        }
        if (planetshine) {
            R visibility = air_light_visibility(c, camera_ray(p, reflected_direction), reflector);
            if (visibility > 0.0) {
                R light_mie;
                R light_column = density_integral(c, p, reflected_direction, c->sun_steps, &light_mie);
                v3 light_t = expV(neg(air_depth(c, light_column, light_mie)));
                v3 reflected_phase = add(scl(br_step, reflected_rayleigh), Vs(bm_step * reflected_mie));
                inscatter = add(inscatter, mul(mul(mul(mul(t_view, segment), reflected_phase), reflected),
                                               scl(light_t, visibility)));
            }
        }
        t_view = mul(t_view, step_t);
    }
    return inscatter;
}
static v3 camera_air_transmittance(const Ctx *c, v3 origin, v3 dir, R limit, uint32_t seed) {
    R exit = shell_exit(c, origin, dir);
    R t_max = mn(limit, exit);
    if (t_max <= 0.0) return Vs(1.0);
    R H = scale_height(c);
    uint32_t steps = (uint32_t)ceil(t_max / (0.25 * H));
    if (steps < 2u) steps = 2u;
    if (steps > c->view_steps) steps = c->view_steps;
    R dt = t_max / (R)steps;
    uint32_t jitter_rng = orc_hash(seed ^ 0x3c6ef372u);
    R jd = random_(&jitter_rng);
    R jitter = (steps < c->view_steps || limit < exit) ? jd : 0.5;
    v3 transmittance = Vs(1.0);
    for (uint32_t i = 0; i < steps; ++i) {
        v3 p = add(origin, scl(dir, ((R)i + jitter) * dt));
        R altitude = mx(len(sub(p, atmo_c(c))) - host_radius(c), 0.0);
        R column = exp(-altitude / H) * dt;
        R column_m = two_heights(c) ? exp(-altitude / c->mie_height) * dt : 0.0;
        transmittance = mul(transmittance, expV(neg(air_depth(c, column, column_m))));
    }
    return transmittance;
}

/* --- sky cache (src/shaders/forest.wgsl:3-54) --- */
static v3 sky_texel(const Ctx *c, int x, int y) {
    int yy = y < 0 ? 0 : (y > 7 ? 7 : y);
    const float *t = c->sky + ((uint32_t)yy * 16u + (uint32_t)((x + 16) % 16)) * 4u;
    return V(t[0], t[1], t[2]);
}
static v3 surface_sky_lookup(const Ctx *c, v3 dir) {
    v3 p = to_ground(c, dir);
    if (p.y <= 0.0) return Vs(0.0);
    R u = fmod(atan2(p.z, p.x) / TAU + 1.0, 1.0) * 16.0 - 0.5;
    R v = p.y * 8.0 - 0.5;
    int cx = (int)floor(u), cy = (int)floor(v);
    R fx = fract(u), fy = fract(v);
    return mixV(mixV(sky_texel(c, cx, cy), sky_texel(c, cx + 1, cy), fx),
                mixV(sky_texel(c, cx, cy + 1), sky_texel(c, cx + 1, cy + 1), fx), fy);
}
static int sky_blocked(const Ctx *c, Ray ray) {
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        if (c->s->ground_north.v[3] == (float)(i + 1u)) continue;
        if (sphere_hit(c, ray, i).index != MISS) return 1;
    }
    return landscape_query(c, ray, 1.0e30, 1).index != MISS;
}
static v3 surface_sky(const Ctx *c, Ray origin, v3 normal, uint32_t *rng) {
    R u = random_(rng);
    R phi = TAU * random_(rng);
    v3 dir = basis_direction(normal, sqrt(u) * cos(phi), sqrt(u) * sin(phi), sqrt(1.0 - u));
    v3 light = surface_sky_lookup(c, dir);
    if (maxc(light) <= 0.0) return Vs(0.0);
    Ray ray = {origin.offset, dir, origin.anchor};
    if (sky_blocked(c, ray)) return Vs(0.0);
    R transmission = 1.0;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        RingHit hit = ring_hit(c, ray, i);
        if (hit.index != MISS) transmission *= ring_transmission(c, ray, hit);
    }
    return scl(light, transmission);
}

/* --- ringshine, point discs (src/shaders/pathtrace.wgsl:1237-1345) --- */
static R ring_area(const Ctx *c, uint32_t index) {
    const Body *body = B(c, index);
    R r = body->center.v[3];
    R o = body->ring_params.v[0], i = body->ring_plane.v[3];
    return PI * r * r * (o * o - i * i);
}
static v3 ringshine(const Ctx *c, uint32_t index, v3 origin, v3 normal, int last_vertex, uint32_t *rng) {
    const Body *body = B(c, index);
    if (body->ring_params.v[0] <= body->ring_plane.v[3] || body->ring_params.v[1] <= 0.0f) return Vs(0.0);
    v3 n = V4(body->ring_plane);
    R inner = (R)body->ring_plane.v[3] * (R)body->ring_plane.v[3];
    R outer = (R)body->ring_params.v[0] * (R)body->ring_params.v[0];
    R radius = sqrt(mixR(inner, outer, random_(rng))) * (R)body->center.v[3];
    R phi = TAU * random_(rng);
    v3 point = scl(basis_direction(n, cos(phi), sin(phi), 0.0), radius);
    v3 offset = sub(point, origin);
    R distance2 = dot(offset, offset);
    R distance = sqrt(distance2);
    v3 direction = dvs(offset, distance);
    R cos_surface = dot(normal, direction);
    R cos_ring = fabs(dot(direction, n));
    if (cos_surface <= 0.0 || cos_ring < 1.0e-6) return Vs(0.0);
    Ray br = {origin, direction, index};
    Hit blocker = sphere_hit(c, br, index);
    if (blocker.index != MISS && blocker.distance < distance) return Vs(0.0);
    v3 col;
    R tau = ring_optics(c, index, point, &col);
    if (tau == 0.0) return Vs(0.0);
    R mu_out = dot(neg(direction), n);
    v3 shine = Vs(0.0);
    for (uint32_t slot = 0; slot < light_count(c); ++slot) {
        uint32_t i = light_index(c, slot);
        if (B(c, i)->material.v[3] < 0.5f) continue;
        Ray lr = {point, n, index};
        LightSample light = sample_light(c, lr, i, rng);
        R mu_in = dot(light.direction, n);
        RingHit rh = {0.0, index, point};
        R visibility = light_visibility(c, ring_origin(c, rh, light.direction), i);
        shine = add(shine, scl(mul(ring_bsdf(col, tau, mu_out, mu_in), V4(B(c, i)->material)),
                               fabs(mu_in) * visibility / light.pdf));
    }
    R pdf = distance2 / (ring_area(c, index) * cos_ring);
    R weight = power_weight(pdf, cos_surface / PI);
    if (last_vertex) weight = 1.0;
    return scl(shine, cos_surface * weight / (PI * pdf));
}
static R pixel_angle(const Ctx *c) { return 2.0 * (R)c->s->cam_forward.v[3] / mx((R)c->s->viewport.v[1], 1.0); }
static Hit point_disc_hit(const Ctx *c, Ray ray, Hit closest, R *scale, uint32_t *rng) {
    Hit result = closest;
    R total = 0.0;
    R pixel = pixel_angle(c);
    R angle = clampR(POINT_DISC_RADIUS, 0.7 * pixel, STAR_DOT_PIXELS * pixel);
    uint32_t host = (uint32_t)c->s->atmo_params.v[3] - 1u;
    for (uint32_t i = 0; i < c->s->counts[0]; ++i) {
        if (i == host || c->s->ground_north.v[3] == (float)(i + 1u)) continue;
        R radius = B(c, i)->center.v[3];
        R drawn = angle * len(relative_center(c, ray, i));
        if (radius <= 0.0 || drawn <= radius) continue;
        Hit hit = sphere_hit_radius(c, ray, i, drawn);
        if (hit.distance < closest.distance) {
            R ratio = (radius / drawn) * (radius / drawn);
            total += ratio;
            if (random_(rng) * total < ratio) result = hit;
        }
    }
    if (total > 0.0) *scale = total;
    return result;
}
static R point_gain(const Ctx *c, R extent, R distance) {
    R pixels = 2.0 * extent / mx(distance, 1.0e-30) / pixel_angle(c);
    return mixR(c->s->background[3], 1.0, smooth(1.0, 3.0, pixels));
}

typedef struct { v3 radiance; uint32_t rng; } CameraAir;

/* src/shaders/pathtrace.wgsl:1347-1617 */
static void trace(const Ctx *c, Ray initial, uint32_t *rng, CameraAir primary_air, R out[4], v3 *backdrop_out) {
    const Settings *s = c->s;
    Ray ray = initial;
    v3 throughput = Vs(1.0), radiance = Vs(0.0), backdrop = Vs(0.0);
    R previous_pdf = 0.0;
    Ray previous_vertex = initial;
    uint32_t depth = 0, null_crossings = 0;
    int camera_segment = 1, in_air = 0, sky_on_escape = 0, point_gained = 0;
    R ring_cap = 1.0e30;
    v3 ring_base = Vs(0.0);
    uint32_t shine_body = MISS;
    COUNT(EV_PATHS);
    for (;;) {
        COUNT(EV_SEGMENTS);
        Hit hit = closest_hit(c, ray);
        R disc_scale = 1.0;
        if (depth == 0u && null_crossings == 0u && s->background[3] > 0.0f) hit = point_disc_hit(c, ray, hit, &disc_scale, rng);
        RingHit ring = closest_ring(c, ray, hit.distance);
        if (in_air) {
            in_air = 0;
            uint32_t host = (uint32_t)s->atmo_params.v[3] - 1u;
            if (ring.index != MISS || (hit.index < LOCAL_BASE && hit.index != host))
                throughput = mul(throughput, air_transmittance(c, ray, ray.direction));
        }
        if (camera_segment) {
            camera_segment = 0;
            R limit = 1.0e30;
            if (hit.index != MISS) limit = hit.distance;
            if (ring.index != MISS) limit = mn(limit, ring.distance);
            if (s->atmo_params.v[3] >= 0.5f) {
                stage = 1;
                v3 transmittance = camera_air_transmittance(c, ray.offset, ray.direction, limit, *rng);
                stage = 0;
                radiance = add(radiance, mul(throughput, primary_air.radiance));
                throughput = mul(throughput, transmittance);
                *rng = primary_air.rng;
            }
        }
        if (ring.index != MISS) {
            const Body *planet = B(c, ring.index);
            if (depth == 0u && !point_gained && s->background[3] > 0.0f) {
                point_gained = 1;
                throughput = scl(throughput, point_gain(c, (R)planet->center.v[3] * (R)planet->ring_params.v[0], ring.distance));
            }
            v3 n = V4(planet->ring_plane);
            v3 col;
            R tau = ring_optics(c, ring.index, ring.point, &col);
            R mu_out = dot(neg(ray.direction), n);
            R transmission = ring_transmission(c, ray, ring);
            R scatter_probability = one_minus_exp(tau / mx(fabs(mu_out), 1.0e-7));
            R bsdf_pdf = scatter_probability / (4.0 * PI);
            int last_vertex = depth + 1u == s->counts[3];
            R ring_mis = 1.0;
            if (shine_body == ring.index && null_crossings == 0u) {
                R cos_ring = mx(fabs(dot(ray.direction, n)), 1.0e-6);
                R shine_pdf = ring.distance * ring.distance / (ring_area(c, ring.index) * cos_ring);
                ring_mis = power_weight(previous_pdf, shine_pdf);
            }
            shine_body = MISS;
            R unshadowed = 0.0, direct = 0.0;
            for (uint32_t slot = 0; slot < light_count(c); ++slot) {
                uint32_t i = light_index(c, slot);
                if (B(c, i)->material.v[3] < 0.5f || tau == 0.0) continue;
                Ray origin = {ring.point, n, ring.index};
                LightSample light = sample_light(c, origin, i, rng);
                R mu_in = dot(light.direction, n);
                R visibility = light_visibility(c, ring_origin(c, ring, light.direction), i);
                R weight = power_weight(light.pdf, bsdf_pdf);
                if (last_vertex) weight = 1.0;
                v3 lit = scl(mul(mul(throughput, ring_bsdf(col, tau, mu_out, mu_in)), V4(B(c, i)->material)),
                             fabs(mu_in) / light.pdf);
                radiance = add(radiance, scl(lit, visibility * weight * ring_mis));
                unshadowed += lum(lit);
                direct += lum(lit) * visibility;
            }
            if (random_(rng) < transmission) {
                ray = ring_origin(c, ring, ray.direction);
                null_crossings += 1u;
                if (null_crossings > s->counts[0]) break;
                continue;
            }
            if (last_vertex || scatter_probability == 0.0) break;
            if (ring_cap > 1.0e29) {
                ring_cap = mx(direct, RING_INDIRECT_CAP * unshadowed);
                ring_base = radiance;
            }
            R z = 1.0 - 2.0 * random_(rng);
            R phi = TAU * random_(rng);
            R radial = sqrt(mx(0.0, 1.0 - z * z));
            v3 direction = basis_direction(n, radial * cos(phi), radial * sin(phi), z);
            throughput = mul(throughput, scl(ring_bsdf(col, tau, mu_out, z), fabs(z) / bsdf_pdf));
            previous_pdf = bsdf_pdf;
            sky_on_escape = 0;
            ray = ring_origin(c, ring, direction);
            previous_vertex = ray;
            depth += 1u;
            null_crossings = 0u;
            if (depth >= 3u) {
                R survival = clampR(maxc(throughput), 0.05, 0.95);
                if (random_(rng) >= survival) break;
                throughput = dvs(throughput, survival);
            }
            continue;
        }
        if (hit.index == MISS) {
            radiance = add(radiance, mul(throughput, environment(c, ray.direction, depth == 0u)));
            if (sky_on_escape) radiance = add(radiance, mul(throughput, surface_sky_lookup(c, ray.direction)));
            if (depth == 0u) {
                backdrop = mul(milky_way(c, ray.direction), throughput);
                radiance = add(radiance, backdrop);
            }
            break;
        }
        int local = hit.index >= LOCAL_BASE;
        const Body *body = local ? 0 : B(c, hit.index);
        throughput = scl(throughput, disc_scale);
        if (!local && body->material.v[3] > 0.5f) {
            R weight = 1.0;
            if (depth > 0u) {
                R light_pdf = 1.0 / (TAU * cone_width(c, previous_vertex, hit.index));
                weight = power_weight(previous_pdf, light_pdf);
            }
            radiance = add(radiance, scl(mul(throughput, V4(body->material)), weight));
            break;
        }
        if (!local && depth == 0u && !point_gained && s->background[3] > 0.0f) {
            point_gained = 1;
            R extent = (R)body->center.v[3] * mx(1.0, (R)body->ring_params.v[0]);
            throughput = scl(throughput, point_gain(c, extent, hit.distance));
        }
        v3 normal = hit.normal;
        if (local) normal = landscape_shading_normal(c, ray, hit);
        Ray outgoing;
        if (local) outgoing = landscape_outgoing(c, ray, hit);
        else {
            Ray o = {scl(normal, (R)body->center.v[3] * (1.0 + 2.0e-6)), normal, hit.index};
            outgoing = o;
        }
        COUNT(EV_VERTEX);
        v3 direct_light = Vs(0.0);
        int last_vertex = depth + 1u == s->counts[3];
        for (uint32_t slot = 0; slot < light_count(c); ++slot) {
            uint32_t i = light_index(c, slot);
            if (B(c, i)->material.v[3] < 0.5f) continue;
            LightSample sample = sample_light(c, outgoing, i, rng);
            R cosine = mx(dot(normal, sample.direction), 0.0);
            if (cosine <= 0.0) continue;
            outgoing.direction = sample.direction;
            R visibility = light_visibility(c, outgoing, i);
            if (visibility > 0.0) {
                R bsdf_pdf = cosine / PI;
                R weight = power_weight(sample.pdf, bsdf_pdf);
                if (last_vertex) weight = 1.0;
                direct_light = add(direct_light, scl(mul(V4(B(c, i)->material), air_transmittance(c, outgoing, sample.direction)),
                                                     bsdf_pdf * weight * visibility / sample.pdf));
            }
        }
        v3 albedo = local ? landscape_albedo(c, ray, hit) : surface_albedo(body, normal);
        radiance = add(radiance, mul(mul(throughput, albedo), direct_light));
        shine_body = MISS;
        if (!local) {
            radiance = add(radiance, mul(mul(throughput, albedo), ringshine(c, hit.index, outgoing.offset, normal, last_vertex, rng)));
            shine_body = hit.index;
        }
        int sky_lit = local && s->ground_counts[1] != 0u && s->atmo_params.v[3] > 0.5f;
        if (sky_lit && last_vertex) radiance = add(radiance, mul(mul(throughput, albedo), surface_sky(c, outgoing, normal, rng)));
        if (last_vertex) break;
        R u = random_(rng);
        R phi = TAU * random_(rng);
        R cosine = sqrt(1.0 - u);
        outgoing.direction = basis_direction(normal, sqrt(u) * cos(phi), sqrt(u) * sin(phi), cosine);
        previous_pdf = cosine / PI;
        throughput = mul(throughput, albedo);
        if (depth >= 2u) {
            R survival = clampR(maxc(throughput), 0.05, 0.95);
            if (random_(rng) >= survival) break;
            throughput = dvs(throughput, survival);
        }
        ray = outgoing;
        previous_vertex = outgoing;
        in_air = inside_air(c, outgoing);
        sky_on_escape = sky_lit;
        depth += 1u;
        null_crossings = 0u;
    }
    if (ring_cap < 1.0e29) {
        v3 indirect = sub(radiance, ring_base);
        R l = lum(indirect);
        if (l > ring_cap) radiance = add(ring_base, scl(indirect, ring_cap / l));
    }
    out[0] = radiance.x;
    out[1] = radiance.y;
    out[2] = radiance.z;
    out[3] = lum(backdrop);
    *backdrop_out = backdrop;
}

/* src/shaders/pathtrace.wgsl:1619-1626 */
static v3 camera_direction(const Ctx *c, uint32_t x, uint32_t y, uint32_t *rng) {
    R jx = random_(rng);
    R jy = random_(rng);
    R u = ((R)x + jx) / (R)c->s->viewport.v[0];
    R v = ((R)y + jy) / (R)c->s->viewport.v[1];
    R ndc_x = 2.0 * u - 1.0, ndc_y = 1.0 - 2.0 * v;
    R tangent = c->s->cam_forward.v[3];
    return nrm(V(ndc_x * tangent * (R)c->s->viewport.v[3], ndc_y * tangent, -1.0));
}

/* integrate_camera_air (src/shaders/pathtrace.wgsl:1630-1643) */
static CameraAir integrate_camera_air(const Ctx *c, uint32_t x, uint32_t y, uint32_t sequence) {
    uint32_t width = (uint32_t)c->s->viewport.v[0];
    uint32_t pixel = y * width + x;
    uint32_t rng = orc_hash(pixel ^ orc_hash(sequence + 0x9e3779b9u));
    Ray ray = camera_ray(Vs(0.0), camera_direction(c, x, y, &rng));
    Hit hit = closest_hit(c, ray);
    R disc_scale = 1.0;
    if (c->s->background[3] > 0.0f) hit = point_disc_hit(c, ray, hit, &disc_scale, &rng);
    RingHit ring = closest_ring(c, ray, hit.distance);
    CameraAir air;
    stage = 1;
    air.radiance = atmosphere(c, ray.offset, ray.direction, mn(hit.distance, ring.distance), &rng);
    stage = 0;
    air.rng = rng;
    return air;
}

static Ctx make_ctx(const orc_scene *sc) {
    Ctx c;
    c.s = (const Settings *)sc->settings;
    c.b = (const Body *)sc->bodies;
    c.objects = (const Prim *)sc->objects;
    c.sky = sc->sky_light;
    c.terrain = sc->terrain;
    c.catalogue = (const RayStar *)sc->catalogue;
    c.ncat = sc->catalogue_count;
    c.cells = sc->sky_cells;
    c.milky = sc->milky_way;
    c.ring_profile = sc->ring_profile;
    c.ring_profile_body = sc->ring_profile ? sc->ring_profile_body : MISS;
    c.view_steps = sc->air_view_steps ? sc->air_view_steps : ATMO_VIEW_STEPS;
    c.sun_steps = sc->air_sun_steps ? sc->air_sun_steps : ATMO_SUN_STEPS;
    c.mie_height = sc->mie_height;
    c.accumulated = ((const Settings *)sc->settings)->counts[1];
    return c;
}

/*
 * One `encode` pass over rows [row_begin, row_end): the reference's main
 * (src/shaders/pathtrace.wgsl:1645-1675) for counts.z samples, or, with air
 * on, counts.z per-sample (integrate_camera_air, main) dispatch pairs with
 * counts.z = 1 (src/pathtracer.rs:861-882,971-976). `accumulation` and
 * `backdrop` are W*H vec4; `moments` (optional) receives the running mean of
 * squared radiance (rgb) and the sample count (w).
 */
void orc_encode_rows(const orc_scene *sc, uint32_t row_begin, uint32_t row_end, float *accumulation, float *backdrop,
                     float *moments) {
    Ctx c0 = make_ctx(sc);
    const Settings *s = c0.s;
    const uint32_t width = (uint32_t)s->viewport.v[0];
    const uint32_t count = s->counts[2];
    const uint32_t old_count = s->counts[1];
    const int has_air = s->atmo_params.v[3] >= 0.5f;
    const uint32_t base_sequence = s->sampling[0];
#pragma omp parallel
    {
#pragma omp for schedule(dynamic, 1)
    for (uint32_t y = row_begin; y < row_end; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            Ctx c = c0;
            uint32_t pixel = y * width + x;
            float *acc = accumulation + (size_t)pixel * 4u;
            float *bd = backdrop + (size_t)pixel * 4u;
            float *mo = moments ? moments + (size_t)pixel * 4u : 0;
            if (!has_air) {
                R sum[4] = {0, 0, 0, 0}, bsum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
                CameraAir none = {{0, 0, 0}, 0};
                for (uint32_t k = 0; k < count; ++k) {
                    uint32_t rng = orc_hash(pixel ^ orc_hash(base_sequence + k + 0x9e3779b9u));
                    v3 dir = camera_direction(&c, x, y, &rng);
                    R v[4];
                    v3 b;
                    trace(&c, camera_ray(Vs(0.0), dir), &rng, none, v, &b);
                    for (int i = 0; i < 4; ++i) sum[i] += v[i];
                    bsum[0] += b.x; bsum[1] += b.y; bsum[2] += b.z;
                    for (int i = 0; i < 3; ++i) sq[i] += v[i] * v[i];
                }
                R n = (R)count;
                R mean[4], bmean[3], m2[3];
                for (int i = 0; i < 4; ++i) mean[i] = sum[i] / n;
                for (int i = 0; i < 3; ++i) { bmean[i] = bsum[i] / n; m2[i] = sq[i] / n; }
                uint32_t new_count = old_count + count;
                if (old_count > 0u) {
                    R w = n / (R)new_count;
                    for (int i = 0; i < 4; ++i) mean[i] = (R)acc[i] + (mean[i] - (R)acc[i]) * w;
                    for (int i = 0; i < 3; ++i) bmean[i] = (R)bd[i] + (bmean[i] - (R)bd[i]) * w;
                    if (mo) for (int i = 0; i < 3; ++i) m2[i] = (R)mo[i] + (m2[i] - (R)mo[i]) * w;
                }
                for (int i = 0; i < 4; ++i) acc[i] = (float)mean[i];
                for (int i = 0; i < 3; ++i) bd[i] = (float)bmean[i];
                bd[3] = 0.0f;
                if (mo) { for (int i = 0; i < 3; ++i) mo[i] = (float)m2[i]; mo[3] = (float)new_count; }
            } else {
                for (uint32_t k = 0; k < count; ++k) {
                    c.accumulated = old_count + k;
                    uint32_t sequence = base_sequence + k;
                    CameraAir air = integrate_camera_air(&c, x, y, sequence);
                    uint32_t rng = orc_hash(pixel ^ orc_hash(sequence + 0x9e3779b9u));
                    v3 dir = camera_direction(&c, x, y, &rng);
                    R v[4];
                    v3 b;
                    trace(&c, camera_ray(Vs(0.0), dir), &rng, air, v, &b);
                    R bb[3] = {b.x, b.y, b.z};
                    uint32_t oc = old_count + k;
                    if (oc > 0u) {
                        R w = 1.0 / (R)(oc + 1u);
                        for (int i = 0; i < 4; ++i) acc[i] = (float)((R)acc[i] + (v[i] - (R)acc[i]) * w);
                        for (int i = 0; i < 3; ++i) bd[i] = (float)((R)bd[i] + (bb[i] - (R)bd[i]) * w);
                        if (mo) for (int i = 0; i < 3; ++i) mo[i] = (float)((R)mo[i] + (v[i] * v[i] - (R)mo[i]) * w);
                    } else {
                        for (int i = 0; i < 4; ++i) acc[i] = (float)v[i];
                        for (int i = 0; i < 3; ++i) bd[i] = (float)bb[i];
                        if (mo) for (int i = 0; i < 3; ++i) mo[i] = (float)(v[i] * v[i]);
                    }
                    bd[3] = 0.0f;
                    if (mo) mo[3] = (float)(oc + 1u);
                }
            }
        }
    }
    flush_counters();
    }
}

/* cache_terrain_geometry (src/shaders/terrain.wgsl:242-285) */
void orc_terrain_cache(const orc_scene *sc) {
    Ctx c = make_ctx(sc);
    TL t = terrain_layout(&c);
    const R scale = terrain_scale(&c);
#pragma omp parallel for schedule(static)
    for (uint64_t index = 0; index < t.vertices; ++index) {
        if (index < t.nodes) {
            uint32_t node = (uint32_t)index;
            uint32_t level = first_leading_bit(3u * node + 1u) / 2u;
            uint32_t size = t.cells >> level;
            uint32_t morton = node - ((1u << (2u * level)) - 1u) / 3u;
            uint32_t nx = compact_bits(morton) * size, nz = compact_bits(morton >> 1u) * size;
            R lox = terrain_coordinate(&c, &t, nx), loz = terrain_coordinate(&c, &t, nz);
            R hix = terrain_coordinate(&c, &t, nx + size), hiz = terrain_coordinate(&c, &t, nz + size);
            R nearx = clampR(0.0, lox, hix), nearz = clampR(0.0, loz, hiz);
            R farx = mx(fabs(lox), fabs(hix)), farz = mx(fabs(loz), fabs(hiz));
            R lower = (R)c.terrain[t.bounds + (uint64_t)node * 2u] * scale + terrain_curvature(&c, farx, farz) - 0.02;
            R upper = (R)c.terrain[t.bounds + (uint64_t)node * 2u + 1u] * scale + terrain_curvature(&c, nearx, nearz) + 0.02;
            float *q = c.terrain + t.cache + (uint64_t)node * 6u;
            q[0] = (float)lox; q[1] = (float)lower; q[2] = (float)loz;
            q[3] = (float)hix; q[4] = (float)upper; q[5] = (float)hiz;
        }
        uint32_t x = (uint32_t)(index % t.width), z = (uint32_t)(index / t.width);
        R px = terrain_coordinate(&c, &t, x), pz = terrain_coordinate(&c, &t, z);
        R vy = (R)c.terrain[index] * scale + terrain_curvature(&c, px, pz);
        R r = c.s->ground_up.v[3];
        R sq = sqrt(mx(r * r - (px * px + pz * pz), 1.0));
        R cgx = px / sq, cgz = pz / sq;
        v3 normal;
        if (t.compact) {
            uint32_t lx = x > 0u ? x - 1u : 0u, lz = z > 0u ? z - 1u : 0u;
            uint32_t hx = x + 1u < t.cells ? x + 1u : t.cells, hz = z + 1u < t.cells ? z + 1u : t.cells;
            R dx = ((R)c.terrain[(uint64_t)z * t.width + hx] - (R)c.terrain[(uint64_t)z * t.width + lx]) /
                   ((terrain_coordinate(&c, &t, hx) - terrain_coordinate(&c, &t, lx)) / scale);
            R dz = ((R)c.terrain[(uint64_t)hz * t.width + x] - (R)c.terrain[(uint64_t)lz * t.width + x]) /
                   ((terrain_coordinate(&c, &t, hz) - terrain_coordinate(&c, &t, lz)) / scale);
            normal = nrm(V(-dx + cgx, 1.0, -dz + cgz));
        } else {
            normal = nrm(V(-(R)c.terrain[t.vertices + index] + cgx, 1.0, -(R)c.terrain[2u * t.vertices + index] + cgz));
        }
        float *v = c.terrain + t.vertex_cache + index * t.stride;
        if (t.compact) {
            uint32_t qx = (uint32_t)rint(clampR(normal.x * 0.5 + 0.5, 0.0, 1.0) * 4095.0);
            uint32_t qz = (uint32_t)rint(clampR(normal.z * 0.5 + 0.5, 0.0, 1.0) * 4095.0);
            v[0] = (float)vy;
            v[1] = (float)(qx | (qz << 12u));
            continue;
        }
        v[0] = (float)px; v[1] = (float)vy; v[2] = (float)pz;
        v[3] = (float)normal.x; v[4] = (float)normal.y; v[5] = (float)normal.z;
    }
}

/* cache_surface_sky (src/shaders/forest.wgsl:3-17); writes 128 x vec4 */
void orc_sky_cache(const orc_scene *sc, float *sky) {
    Ctx c = make_ctx(sc);
    for (uint32_t index = 0; index < 128u; ++index) {
        uint32_t x = index % 16u, y = index / 16u;
        v3 light = Vs(0.0);
        if (c.s->atmo_params.v[3] > 0.5f && c.s->ground_counts[1] != 0u) {
            R mu = ((R)y + 0.5) / 8.0;
            R phi = TAU * ((R)x + 0.5) / 16.0;
            R r = sqrt(1.0 - mu * mu);
            v3 dir = from_ground(&c, V(r * cos(phi), mu, r * sin(phi)));
            uint32_t rng = orc_hash(index ^ 0x91ce452du);
            light = atmosphere(&c, Vs(0.0), dir, 1.0e30, &rng);
        }
        sky[index * 4u + 0] = (float)light.x;
        sky[index * 4u + 1] = (float)light.y;
        sky[index * 4u + 2] = (float)light.z;
        sky[index * 4u + 3] = 0.0f;
    }
}

/* Probe kinds match sg_probe_kind in include/stargaze_amd.h. */
void orc_probe(const orc_scene *sc, uint32_t kind, const orc_probe_ray *rays, uint32_t count, orc_probe_result *out) {
    Ctx c = make_ctx(sc);
    for (uint32_t i = 0; i < count; ++i) {
        const orc_probe_ray *pr = &rays[i];
        Ray ray = {V(pr->offset[0], pr->offset[1], pr->offset[2]), V(pr->direction[0], pr->direction[1], pr->direction[2]),
                   pr->anchor};
        orc_probe_result r;
        memset(&r, 0, sizeof(r));
        r.index = MISS;
        r.distance = 1.0e30;
        if (kind == 0u || kind == 1u || kind == 5u) {
            Hit h = kind == 0u ? sphere_hit(&c, ray, pr->target) : (kind == 1u ? closest_hit(&c, ray) : landscape_hit(&c, ray));
            r.distance = h.distance;
            r.index = h.index;
            r.normal[0] = h.normal.x; r.normal[1] = h.normal.y; r.normal[2] = h.normal.z;
        } else if (kind == 2u) {
            RingHit h = ring_hit(&c, ray, pr->target);
            r.distance = h.distance;
            r.index = h.index;
            r.extra[0] = h.point.x; r.extra[1] = h.point.y; r.extra[2] = h.point.z;
            if (h.index != MISS) {
                r.normal[0] = ring_transmission(&c, ray, h);
                r.normal[1] = ring_optics(&c, h.index, h.point, 0);
            }
        } else if (kind == 3u) {
            r.distance = light_visibility(&c, ray, pr->target);
            r.index = 0;
        } else if (kind == 4u) {
            v3 n = Vs(0.0);
            R t = terrain_query(&c, ray.offset, ray.direction, 1.0e30, pr->target != 0u, &n);
            r.distance = t;
            r.index = t < 1.0e29 ? TERRAIN_ID : MISS;
            r.normal[0] = n.x; r.normal[1] = n.y; r.normal[2] = n.z;
        } else if (kind == 6u) {
            r.distance = cone_width(&c, ray, pr->anchor == MISS ? 0u : pr->anchor);
        } else if (kind == 7u) {
            /* material probe of src/surface_tests.rs:106-125: the raw style
             * albedo at the local point q = direction (plain style: the
             * clamped surface albedo) */
            const Body *b = B(&c, pr->target);
            uint32_t style = (uint32_t)((R)b->axis_x.v[3] + 0.5);
            v3 a = style > 0u ? style_albedo(b, ray.direction) : surface_albedo(b, ray.direction);
            r.index = 0;
            r.normal[0] = a.x; r.normal[1] = a.y; r.normal[2] = a.z;
        } else if (kind == 9u) {
            /* value noise, its hash and fbm at q = direction (pathtrace.wgsl:338-365) */
            r.index = 0;
            r.distance = vnoise(ray.direction);
            r.normal[0] = hash31(ray.direction);
            r.normal[1] = fbm(ray.direction, pr->target);
        } else if (kind == 10u) {
            /* surface_sky over the last encode's sky cache, and one texel of it
             * (forest.wgsl:3-54) */
            uint32_t rng = orc_hash(pr->target + 817u);
            v3 light = surface_sky(&c, ray, ray.direction, &rng);
            v3 texel = sky_texel(&c, (int)(pr->target % 16u), (int)((pr->target / 16u) % 8u));
            r.index = 0;
            r.normal[0] = light.x; r.normal[1] = light.y; r.normal[2] = light.z;
            r.extra[0] = texel.x; r.extra[1] = texel.y; r.extra[2] = texel.z;
        } else if (kind == 8u) {
            /* ground material (ground.wgsl:597-648) of landscape element
             * anchor at surface-frame point offset (metres) with surface-frame
             * geometric normal direction, seen from target's float bits metres */
            float distance;
            memcpy(&distance, &pr->target, 4);
            v3 p = V(pr->offset[0], pr->offset[1], pr->offset[2]);
            v3 n = from_ground(&c, ray.direction);
            v3 shading = ground_shading_normal(&c, p, (R)distance, n, pr->anchor);
            v3 a = ground_albedo(&c, p, (R)distance, ray.direction, pr->anchor);
            r.index = pr->anchor;
            r.distance = distance;
            r.normal[0] = shading.x; r.normal[1] = shading.y; r.normal[2] = shading.z;
            r.extra[0] = a.x; r.extra[1] = a.y; r.extra[2] = a.z;
        }
        out[i] = r;
    }
}

/* Ring energy probe (src/ring_tests.rs:351-384): the transmitted delta plus
 * the BSDF integrated over incoming cosines, for one outgoing cosine. */
double orc_ring_energy(double albedo, double tau, double mu_out, uint32_t steps) {
    v3 color = Vs((R)albedo);
    R mo = (R)mu_out;
    R energy = exp(-(R)tau / mo);
    for (uint32_t i = 0; i < steps; ++i) {
        R mi = ((R)i + 0.5) / (R)steps;
        v3 f = add(ring_bsdf(color, (R)tau, mo, mi), ring_bsdf(color, (R)tau, mo, -mi));
        energy += f.x * (TAU * mi / (R)steps);
    }
    return (double)energy;
}

/* Size of REAL, so the loader can tell the two builds apart. */
uint32_t orc_real_bytes(void) { return (uint32_t)sizeof(R); }
