// Host build of the product's device helpers, for CPU tests of their logic
// (termination and answers of the stackless traversals). Test code only:
// the product library is stargaze_amd/libstargaze_amd.so.
#include <cstring>
#include <vector>

#include "../../stargaze_amd/csrc/sg_integrator.h"

using namespace sg;

extern "C" {

// Product terrain layout for `settings` (floats of caller data and of the
// whole buffer) and the product's cache pass over `terrain` on the host.
void hc_terrain_layout(const sg_settings *settings, uint64_t *data, uint64_t *total) {
    TerrainLayout t = terrain_layout(settings->globals.ground_counts[2]);
    *data = t.data;
    *total = t.total;
}
void hc_terrain_cache(const sg_settings *settings, float *terrain) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = *settings;
    sc.t = terrain_layout(settings->globals.ground_counts[2]);
    for (uint64_t i = 0; i < sc.t.vertices; ++i) terrain_cache_entry(sc, terrain, i);
}

// Landscape query (kind 5) or terrain query (kind 4) for `count` probe rays.
// `bvh` may be null (flat primitive loop). Terrain may be null when off.
int hc_query(const sg_settings *settings, const sg_sphere *bodies, const sg_primitive *objects,
             const float *terrain, const float *bvh_nodes, const uint32_t *bvh_prims,
             uint32_t kind, const sg_probe_ray *rays, uint32_t count, sg_probe_result *out) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = *settings;
    sc.bodies = reinterpret_cast<const Body *>(bodies);
    std::vector<Centre> centres(settings->counts[0]);
    fill_centres(sc.bodies, settings->counts[0], centres.data());
    sc.centres = centres.data();
    sc.objects = objects;
    sc.terrain = terrain;
    sc.bvh = reinterpret_cast<const float4 *>(bvh_nodes);
    sc.bvh_prims = bvh_prims;
    sc.t = terrain_layout(settings->globals.ground_counts[2]);
    for (uint32_t i = 0; i < count; ++i) {
        const sg_probe_ray &pr = rays[i];
        Ray ray{d3{pr.offset[0], pr.offset[1], pr.offset[2]},
                f3{pr.direction[0], pr.direction[1], pr.direction[2]}, pr.anchor};
        sg_probe_result r;
        std::memset(&r, 0, sizeof(r));
        if (kind == SG_PROBE_TERRAIN) {
            float4 h = terrain_query(sc, tof(ray.offset), ray.direction, 1.0e30f, pr.target != 0u);
            r.distance = h.w;
            r.index = h.w < 1.0e29f ? SG_TERRAIN_ID : SG_MISS;
            r.normal[0] = h.x; r.normal[1] = h.y; r.normal[2] = h.z;
        } else {
            Hit h = landscape_query(sc, ray, 1.0e30f, pr.target != 0u);
            r.distance = h.distance;
            r.index = h.index;
            r.normal[0] = h.normal.x; r.normal[1] = h.normal.y; r.normal[2] = h.normal.z;
        }
        out[i] = r;
    }
    return 0;
}
}

extern "C" {
// One pass of counts.z samples per pixel from `settings` (no landscape, no
// catalogue or sky map): the per-pixel mean radiance (rgb) into `mean`, and,
// when `gradient` is not null, the product adjoint for d(loss)/d(mean) =
// d_image added into `gradient` (GRAD_COUNT floats).
int hc_pass(const sg_settings *settings, const sg_sphere *bodies, const float *ring_profile, uint32_t ring_body,
            const float *density_profile, const float *d_image, double *mean, float *gradient) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = *settings;
    sc.bodies = reinterpret_cast<const Body *>(bodies);
    std::vector<Centre> centres(settings->counts[0]);
    fill_centres(sc.bodies, settings->counts[0], centres.data());
    sc.centres = centres.data();
    sc.t = terrain_layout(settings->globals.ground_counts[2]);
    sc.ring_profile = ring_profile;
    sc.ring_profile_body = ring_profile ? ring_body : SG_MISS;
    sc.density_profile = density_profile;
    sc.air_view_steps = ATMO_VIEW_STEPS;
    sc.air_sun_steps = ATMO_SUN_STEPS;
    float density[GRAD_PRIVATE] = {};
    const uint32_t width = (uint32_t)settings->globals.viewport[0];
    const uint32_t height = (uint32_t)settings->globals.viewport[1];
    const uint32_t count = settings->counts[2];
    const bool has_air = settings->globals.atmo_params[3] >= 0.5f;
    for (uint32_t pixel = 0; pixel < width * height; ++pixel) {
        const uint32_t x = pixel % width, y = pixel / width;
        double sum[3] = {0.0, 0.0, 0.0};
        for (uint32_t s = 0; s < count; ++s) {
            const uint32_t accumulated = has_air ? settings->counts[1] + s : settings->counts[1];
            uint32_t rng = path_seed(pixel, settings->sampling[0] + s);
            f3 direction = camera_direction(sc, x, y, rng);
            PathResult r;
            if (gradient) {
                const float *d = d_image + 3u * pixel;
                r = adjoint_path(sc, direction, rng, accumulated, f3{d[0], d[1], d[2]} * (1.0f / (float)count),
                                 gradient, density, 1u);
            } else {
                r = trace_path(sc, direction, rng, accumulated);
            }
            sum[0] += r.value.x;
            sum[1] += r.value.y;
            sum[2] += r.value.z;
        }
        for (int c = 0; c < 3; ++c) mean[3u * pixel + c] = sum[c] / count;
    }
    return 0;
}
}

extern "C" {
// One pass of the product's fused integrator over a frame with a landscape
// (terrain with the product's cache, prop BVH or flat loop, sky cache filled
// here by the product's texel function): per-pixel mean rgb into `mean`.
// `air` = {view steps, light steps, Mie scale height bits}.
int hc_render(const sg_settings *settings, const sg_sphere *bodies, const sg_primitive *objects, const float *terrain,
              const float *bvh_nodes, const uint32_t *bvh_prims, const uint32_t *air, float *mean) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = *settings;
    sc.bodies = reinterpret_cast<const Body *>(bodies);
    std::vector<Centre> centres(settings->counts[0]);
    fill_centres(sc.bodies, settings->counts[0], centres.data());
    sc.centres = centres.data();
    sc.objects = objects;
    sc.terrain = terrain;
    sc.bvh = reinterpret_cast<const float4 *>(bvh_nodes);
    sc.bvh_prims = bvh_prims;
    sc.t = terrain_layout(settings->globals.ground_counts[2]);
    sc.air_view_steps = air[0];
    sc.air_sun_steps = air[1];
    std::memcpy(&sc.mie_height, &air[2], 4);
    std::vector<float4> sky(SG_SKY_TEXELS);
    sc.sky_light = sky.data();
    for (uint32_t y = 0; y < 8u; ++y)
        for (uint32_t x = 0; x < 16u; ++x) {
            f3 t = sky_cache_texel(sc, x, y);
            sky[y * 16u + x] = make_float4(t.x, t.y, t.z, 0.0f);
        }
    const uint32_t width = (uint32_t)settings->globals.viewport[0];
    const uint32_t height = (uint32_t)settings->globals.viewport[1];
    const uint32_t count = settings->counts[2];
    const bool has_air = settings->globals.atmo_params[3] >= 0.5f;
    for (uint32_t pixel = 0; pixel < width * height; ++pixel) {
        const uint32_t x = pixel % width, y = pixel / width;
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t s = 0; s < count; ++s) {
            const uint32_t accumulated = has_air ? settings->counts[1] + s : settings->counts[1];
            uint32_t rng = path_seed(pixel, settings->sampling[0] + s);
            f3 direction = camera_direction(sc, x, y, rng);
            PathResult r = trace_path(sc, direction, rng, accumulated);
            sum[0] += r.value.x;
            sum[1] += r.value.y;
            sum[2] += r.value.z;
        }
        for (int c = 0; c < 3; ++c) mean[3u * pixel + c] = sum[c] / (float)count;
    }
    return 0;
}
}

#include "../../stargaze_amd/csrc/sg_bvh.h"

extern "C" {
// The library's BVH builder on a primitive list: returns the node count
// (2 float4 per node) and fills up to `capacity` nodes / prims.
uint32_t hc_build_bvh(const sg_primitive *objects, uint32_t n, float *nodes, uint32_t *prims, uint32_t capacity,
                      uint32_t *prim_count) {
    std::vector<float4> nv;
    std::vector<uint32_t> order;
    sg_host::build_bvh(objects, n, nv, order);
    const uint32_t count = (uint32_t)(nv.size() / 2);
    *prim_count = (uint32_t)order.size();
    if (count > capacity || order.size() > capacity) return count;
    std::memcpy(nodes, nv.data(), nv.size() * sizeof(float4));
    std::memcpy(prims, order.data(), order.size() * sizeof(uint32_t));
    return count;
}
}
