// Rings, landscape props, terrain quadtree and closest-hit / visibility
// queries (src/shaders/rings.wgsl, ground.wgsl:1-251,597-666, terrain.wgsl,
// forest.wgsl:56-136 and pathtrace.wgsl:173-185).
#pragma once
#include "sg_device.h"

namespace sg {

// --- local surface frame (src/shaders/ground.wgsl:17-61) -------------------
SG_FN d3 camera_origin(const Scene &sc, const Ray &ray) {
    if (ray.anchor == SG_MISS) return ray.offset;
    return body_centre(sc, ray.anchor) + ray.offset;
}
SG_FN f3 g_east(const Scene &sc) { const float *v = sc.s.globals.ground_east; return f3{v[0], v[1], v[2]}; }
SG_FN f3 g_up(const Scene &sc) { const float *v = sc.s.globals.ground_up; return f3{v[0], v[1], v[2]}; }
SG_FN f3 g_north(const Scene &sc) { const float *v = sc.s.globals.ground_north; return f3{v[0], v[1], v[2]}; }
SG_FN f3 to_ground(const Scene &sc, f3 v) { return f3{dot(v, g_east(sc)), dot(v, g_up(sc)), dot(v, g_north(sc))}; }
SG_FN d3 to_ground_d(const Scene &sc, d3 v) {
    return d3{dot(v, tod(g_east(sc))), dot(v, tod(g_up(sc))), dot(v, tod(g_north(sc)))};
}
SG_FN f3 from_ground(const Scene &sc, f3 v) { return g_east(sc) * v.x + g_up(sc) * v.y + g_north(sc) * v.z; }
SG_FN d3 from_ground_d(const Scene &sc, d3 v) {
    return tod(g_east(sc)) * v.x + tod(g_up(sc)) * v.y + tod(g_north(sc)) * v.z;
}
// Metre-frame origin, FP64 (src/shaders/ground.wgsl:29-32).
SG_FN d3 ground_origin(const Scene &sc, const Ray &ray) {
    d3 o = to_ground_d(sc, camera_origin(sc, ray) * METRES_PER_AU_D);
    o.y += (double)sc.s.globals.ground_east[3];
    return o;
}
SG_FN f3 rotate_north(f3 p, float angle) {
    float c = cosf(angle), s = sinf(angle);
    return f3{c * p.x - s * p.y, s * p.x + c * p.y, p.z};
}
SG_FN f3 rock_angles(float packed) {
    uint32_t bits = (uint32_t)packed;
    return f3{(float)(bits & 255u), (float)((bits >> 8) & 255u), (float)((bits >> 16) & 255u)} * (TAU / 256.0f);
}
SG_FN f3 rotate_east(f3 p, float angle) {
    float c = cosf(angle), s = sinf(angle);
    return f3{p.x, c * p.y - s * p.z, s * p.y + c * p.z};
}
SG_FN f3 rotate_up(f3 p, float angle) {
    float c = cosf(angle), s = sinf(angle);
    return f3{c * p.x + s * p.z, p.y, -s * p.x + c * p.z};
}
SG_FN f3 rock_to_local(f3 p, f3 a) { return rotate_east(rotate_up(rotate_north(p, -a.z), -a.y), -a.x); }
SG_FN f3 rock_from_local(f3 p, f3 a) { return rotate_north(rotate_up(rotate_east(p, a.x), a.y), a.z); }

SG_FN float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// src/shaders/ground.wgsl:64-78
SG_FN float2 box_interval(f3 p, f3 d, f3 extent) {
    float nearv = -1.0e30f, farv = 1.0e30f;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        float pa = comp(p, axis), da = comp(d, axis), ea = comp(extent, axis);
        if (fabsf(da) < 1.0e-12f) {
            if (fabsf(pa) > ea) return make_float2(1.0e30f, -1.0e30f);
        } else {
            float a = (-ea - pa) / da;
            float b = (ea - pa) / da;
            nearv = fmaxf(nearv, fminf(a, b));
            farv = fminf(farv, fmaxf(a, b));
        }
    }
    return make_float2(nearv, farv);
}

// src/shaders/ground.wgsl:82-121: box with 8 clipped corners.
SG_FN float4 rock_hit(f3 p, f3 d, f3 e) {
    f3 origin = p / e, direction = d / e;
    float nearv = -1.0e30f, farv = 1.0e30f;
    f3 near_normal = f3s(0.f), far_normal = f3s(0.f);
    const float4 none = make_float4(0.f, 0.f, 0.f, 1.0e30f);
#pragma unroll
    for (uint32_t face = 0; face < 14u; ++face) {
        f3 n = f3s(0.f);
        float limit = 1.0f;
        if (face < 6u) {
            float sgn = (face & 1u) == 1u ? 1.0f : -1.0f;
            if (face / 2u == 0u) n.x = sgn; else if (face / 2u == 1u) n.y = sgn; else n.z = sgn;
        } else {
            uint32_t bits = face - 6u;
            n = f3{(bits & 1u) ? 1.f : -1.f, (bits & 2u) ? 1.f : -1.f, (bits & 4u) ? 1.f : -1.f};
            limit = 1.65f + 0.08f * (float)(bits % 3u);
        }
        float denominator = dot(n, direction);
        float gap = limit - dot(n, origin);
        if (fabsf(denominator) < 1.0e-12f) {
            if (gap < 0.0f) return none;
        } else {
            float distance = gap / denominator;
            if (denominator < 0.0f && distance > nearv) { nearv = distance; near_normal = n / e; }
            if (denominator > 0.0f && distance < farv) { farv = distance; far_normal = n / e; }
        }
        if (nearv > farv) return none;
    }
    if (farv <= 0.001f) return none;
    f3 n = nearv > 0.001f ? normalize(near_normal) : normalize(far_normal);
    return make_float4(n.x, n.y, n.z, nearv > 0.001f ? nearv : farv);
}

// --- Möller–Trumbore (src/shaders/terrain.wgsl:65-81) ----------------------
SG_FN f3 terrain_triangle(f3 o, f3 d, f3 a, f3 b, f3 c) {
    f3 ab = b - a, ac = c - a;
    f3 p = cross(d, ac);
    float determinant = dot(ab, p);
    if (fabsf(determinant) < 1.0e-12f) return f3{0.f, 0.f, 1.0e30f};
    float inverse = 1.0f / determinant;
    f3 relative = o - a;
    float u = dot(relative, p) * inverse;
    f3 q = cross(relative, ab);
    float v = dot(d, q) * inverse;
    float distance = dot(ac, q) * inverse;
    if (u < -1.0e-6f || v < -1.0e-6f || u + v > 1.000001f || distance <= 0.001f) return f3{0.f, 0.f, 1.0e30f};
    return f3{u, v, distance};
}

// --- vegetation (src/shaders/forest.wgsl:56-136) -----------------------------
SG_FN float4 trunk_hit(f3 p, f3 d, f3 e) {
    f3 o = p / e, v = d / e;
    float4 result = make_float4(0.f, 0.f, 0.f, 1.0e30f);
    float a = v.x * v.x + v.z * v.z;
    float b = o.x * v.x + o.z * v.z;
    float c = o.x * o.x + o.z * o.z - 1.0f;
    float disc = b * b - a * c;
    if (a > 1.0e-15f && disc >= 0.0f) {
        float sq = sqrtf(disc);
        for (uint32_t side = 0; side < 2u; ++side) {
            float t = (-b + (side == 1u ? sq : -sq)) / a;
            f3 q = o + v * t;
            if (t > 0.001f && t < result.w && fabsf(q.y) <= 1.0f) {
                f3 n = normalize(f3{q.x, 0.f, q.z} / e);
                result = make_float4(n.x, n.y, n.z, t);
            }
        }
    }
    if (fabsf(v.y) > 1.0e-15f) {
        for (uint32_t side = 0; side < 2u; ++side) {
            float sign_y = side == 1u ? 1.0f : -1.0f;
            float t = (sign_y - o.y) / v.y;
            f3 q = o + v * t;
            if (t > 0.001f && t < result.w && q.x * q.x + q.z * q.z <= 1.0f) result = make_float4(0.f, sign_y, 0.f, t);
        }
    }
    return result;
}
SG_FN float4 conifer_hit(f3 p, f3 d, f3 e) {
    float4 result = trunk_hit(p + f3{0.f, e.y * 0.08f, 0.f}, d, f3{e.x * 0.065f, e.y * 0.92f, e.z * 0.065f});
    f3 o = p / e, v = d / e;
    for (uint32_t tier = 0; tier < 6u; ++tier) {
        float t = (float)tier;
        float base = -0.6f + t * 0.27f;
        float height = fminf(0.66f - t * 0.045f, 1.0f - base);
        float radius = 0.94f - t * 0.148f;
        float ox = sinf(t * 4.1f) * 0.022f, oz = cosf(t * 3.7f) * 0.022f;
        f3 q = o - f3{ox, base + height, oz};
        float k = radius / height;
        float a = v.x * v.x + v.z * v.z - k * k * v.y * v.y;
        float b = q.x * v.x + q.z * v.z - k * k * q.y * v.y;
        float c = q.x * q.x + q.z * q.z - k * k * q.y * q.y;
        float disc = b * b - a * c;
        if (fabsf(a) < 1.0e-12f || disc < 0.0f) continue;
        float sq = sqrtf(disc);
        for (uint32_t side = 0; side < 2u; ++side) {
            float distance = (-b + (side == 1u ? sq : -sq)) / a;
            f3 h = q + v * distance;
            if (distance <= 0.001f || distance >= result.w || h.y > 0.0f || h.y < -height) continue;
            float angle = atan2f(h.z, h.x);
            float skirt = -h.y / height;
            float branches = 0.5f + 0.5f * sinf(angle * (9.0f + t) + t * 3.7f);
            float needle = vnoise((p + d * distance) * f3{9.0f, 17.0f, 9.0f} + f3s(t * 7.0f));
            if (skirt > 0.76f + 0.19f * branches || (skirt > 0.4f && needle < 0.26f)) continue;
            f3 n = normalize(f3{h.x, -k * k * h.y, h.z} / e);
            result = make_float4(n.x, n.y, n.z, distance);
        }
    }
    return result;
}
SG_FN float4 grass_hit(f3 p, f3 d, f3 e) {
    float4 result = make_float4(0.f, 0.f, 0.f, 1.0e30f);
    for (uint32_t blade = 0; blade < 9u; ++blade) {
        float i = (float)blade;
        float angle = i * 2.39996f;
        f3 side = {cosf(angle), 0.0f, sinf(angle)};
        f3 base = side * (e.x * 0.15f) - f3{0.f, e.y, 0.f};
        f3 tip = side * (e.x * (0.6f + 0.3f * sinf(i * 3.0f))) + f3{0.f, e.y * (0.5f + 0.5f * cosf(i * 1.3f)), 0.f};
        f3 width = f3{-side.z, 0.0f, side.x} * 0.035f;
        f3 hit = terrain_triangle(p, d, base - width, base + width, tip);
        if (hit.z < result.w) {
            f3 normal = normalize(cross(width, tip - base));
            if (dot(normal, d) > 0.0f) normal = -normal;
            result = make_float4(normal.x, normal.y, normal.z, hit.z);
        }
    }
    return result;
}

// --- terrain quadtree (src/shaders/terrain.wgsl:6-234) -----------------------
// Correctly rounded, as the host computes it: grid edges and ray origins
// placed on them must agree to the bit.
SG_FN float terrain_scale(const Scene &sc) { return clampf(div_rn(sc.s.globals.ground_up[3], 200000.0f), 0.0001f, 1.0f); }
SG_FN float terrain_coordinate(const Scene &sc, uint32_t i) {
    // 2 i / cells - 1, exactly: cells is a power of two, so the quotient is
    // a scaling by a power of two (grid edges land on +-1 under any
    // division accuracy).
    float u = ldexpf((float)(2u * i), -(int)__builtin_ctz(sc.t.cells)) - 1.0f;
    return u * fabsf(u) * 16000.0f * terrain_scale(sc);
}
SG_FN float terrain_curvature(const Scene &sc, float px, float pz) {
    float r = sc.s.globals.ground_up[3];
    float d2 = px * px + pz * pz;
    return -d2 / (r + sqrtf(fmaxf(r * r - d2, 0.0f)));
}
SG_FN f3 terrain_vertex(const Scene &sc, uint32_t x, uint32_t z) {
    uint64_t i = sc.t.vertex_cache + ((uint64_t)z * sc.t.width + x) * sc.t.stride;
    if (sc.t.compact) return f3{terrain_coordinate(sc, x), sc.terrain[i], terrain_coordinate(sc, z)};
    return f3{sc.terrain[i], sc.terrain[i + 1], sc.terrain[i + 2]};
}
SG_FN f3 terrain_normal(const Scene &sc, uint32_t x, uint32_t z) {
    uint64_t i = sc.t.vertex_cache + ((uint64_t)z * sc.t.width + x) * sc.t.stride;
    if (sc.t.compact) {
        uint32_t packed = (uint32_t)sc.terrain[i + 1];
        float nx = (float)(packed & 4095u) / 4095.0f * 2.0f - 1.0f;
        float nz = (float)(packed >> 12) / 4095.0f * 2.0f - 1.0f;
        return normalize(f3{nx, sqrtf(fmaxf(1.0f - (nx * nx + nz * nz), 1.0e-6f)), nz});
    }
    return f3{sc.terrain[i + 3], sc.terrain[i + 4], sc.terrain[i + 5]};
}
// Slab test against a cached node AABB (src/shaders/terrain.wgsl:118-130).
SG_FN float2 terrain_interval(const Scene &sc, f3 o, f3 inv, uint32_t node) {
    const float4 *c = reinterpret_cast<const float4 *>(sc.terrain + sc.t.cache) + 2u * (uint64_t)node;
    const float4 l = c[0], h = c[1];
    f3 lo = {l.x, l.y, l.z}, hi = {h.x, h.y, h.z};
    bool px = fabsf(inv.x) >= 1.0e29f, py = fabsf(inv.y) >= 1.0e29f, pz = fabsf(inv.z) >= 1.0e29f;
    if ((px && (o.x < lo.x || o.x > hi.x)) || (py && (o.y < lo.y || o.y > hi.y)) ||
        (pz && (o.z < lo.z || o.z > hi.z)))
        return make_float2(1.0e30f, -1.0e30f);
    f3 a = (lo - o) * inv, b = (hi - o) * inv;
    float ex = px ? -1.0e30f : fminf(a.x, b.x), ey = py ? -1.0e30f : fminf(a.y, b.y),
          ez = pz ? -1.0e30f : fminf(a.z, b.z);
    float xx = px ? 1.0e30f : fmaxf(a.x, b.x), xy = py ? 1.0e30f : fmaxf(a.y, b.y),
          xz = pz ? 1.0e30f : fmaxf(a.z, b.z);
    return make_float2(fmaxf(ex, fmaxf(ey, ez)), fminf(xx, fminf(xy, xz)));
}
// Morton de-interleave (src/shaders/terrain.wgsl:134-140).
SG_FN uint32_t terrain_compact_bits(uint32_t value) {
    uint32_t x = value & 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0f0f0f0fu;
    x = (x | (x >> 4)) & 0x00ff00ffu;
    return (x | (x >> 8)) & 0x0000ffffu;
}
SG_FN uint32_t first_leading_bit(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

// One entry of cache_terrain_geometry (src/shaders/terrain.wgsl:242-285):
// node `index` (if it is a node) gets its curved, 0.02 m-padded AABB and
// vertex `index` its curved position and normal (12-bit packed above 1024
// cells). `terrain` is the scene's terrain buffer.
SG_FN void terrain_cache_entry(const Scene &sc, float *terrain, uint64_t index) {
    const TerrainLayout &t = sc.t;
    if (index >= t.vertices) return;
    const float scale = terrain_scale(sc);
    if (index < t.nodes) {
        uint32_t node = (uint32_t)index;
        uint32_t level = first_leading_bit(3u * node + 1u) / 2u;
        uint32_t size = t.cells >> level;
        uint32_t morton = node - ((1u << (2u * level)) - 1u) / 3u;
        uint32_t nx = terrain_compact_bits(morton) * size, nz = terrain_compact_bits(morton >> 1) * size;
        float lox = terrain_coordinate(sc, nx), loz = terrain_coordinate(sc, nz);
        float hix = terrain_coordinate(sc, nx + size), hiz = terrain_coordinate(sc, nz + size);
        float nearx = clampf(0.0f, lox, hix), nearz = clampf(0.0f, loz, hiz);
        float farx = fmaxf(fabsf(lox), fabsf(hix)), farz = fmaxf(fabsf(loz), fabsf(hiz));
        float lower = terrain[t.bounds + (uint64_t)node * 2u] * scale + terrain_curvature(sc, farx, farz) - 0.02f;
        float upper = terrain[t.bounds + (uint64_t)node * 2u + 1u] * scale + terrain_curvature(sc, nearx, nearz) + 0.02f;
        float4 *c = reinterpret_cast<float4 *>(terrain + t.cache) + 2u * (uint64_t)node;
        c[0] = make_float4(lox, lower, loz, 0.0f);
        c[1] = make_float4(hix, upper, hiz, 0.0f);
    }
    const uint32_t x = (uint32_t)(index % t.width), z = (uint32_t)(index / t.width);
    const float px = terrain_coordinate(sc, x), pz = terrain_coordinate(sc, z);
    const float vy = terrain[index] * scale + terrain_curvature(sc, px, pz);
    const float r = sc.s.globals.ground_up[3];
    const float inv = 1.0f / sqrtf(fmaxf(r * r - (px * px + pz * pz), 1.0f));
    const float cgx = px * inv, cgz = pz * inv;
    f3 normal;
    if (t.compact) {
        uint32_t lx = x > 0u ? x - 1u : 0u, lz = z > 0u ? z - 1u : 0u;
        uint32_t hx = umin(x + 1u, t.cells), hz = umin(z + 1u, t.cells);
        float dx = (terrain[(uint64_t)z * t.width + hx] - terrain[(uint64_t)z * t.width + lx]) /
                   ((terrain_coordinate(sc, hx) - terrain_coordinate(sc, lx)) / scale);
        float dz = (terrain[(uint64_t)hz * t.width + x] - terrain[(uint64_t)lz * t.width + x]) /
                   ((terrain_coordinate(sc, hz) - terrain_coordinate(sc, lz)) / scale);
        normal = normalize(f3{-dx + cgx, 1.0f, -dz + cgz});
    } else {
        normal = normalize(f3{-terrain[t.vertices + index] + cgx, 1.0f, -terrain[2u * t.vertices + index] + cgz});
    }
    float *v = terrain + t.vertex_cache + index * t.stride;
    if (t.compact) {
        uint32_t qx = (uint32_t)rintf(clampf(normal.x * 0.5f + 0.5f, 0.0f, 1.0f) * 4095.0f);
        uint32_t qz = (uint32_t)rintf(clampf(normal.z * 0.5f + 0.5f, 0.0f, 1.0f) * 4095.0f);
        v[0] = vy;
        v[1] = (float)(qx | (qz << 12));
        return;
    }
    v[0] = px; v[1] = vy; v[2] = pz;
    v[3] = normal.x; v[4] = normal.y; v[5] = normal.z;
}

// Near-to-far stack traversal (src/shaders/terrain.wgsl:152-234). Theme 0
// shading: the smooth barycentric normal (terrain_shading is identity there).
SG_FN float4 terrain_query(const Scene &sc, f3 o, f3 d, float initial_limit, bool occlusion) {
    float4 result = make_float4(0.f, 0.f, 0.f, 1.0e30f);
    float limit = initial_limit;
    f3 inv = {(d.x >= 0.f ? 1.f : -1.f) / fmaxf(fabsf(d.x), 1.0e-30f),
              (d.y >= 0.f ? 1.f : -1.f) / fmaxf(fabsf(d.y), 1.0e-30f),
              (d.z >= 0.f ? 1.f : -1.f) / fmaxf(fabsf(d.z), 1.0e-30f)};
    float2 root = terrain_interval(sc, o, inv, 0u);
    SG_COUNT(SG_EV_TERRAIN_BOX);
    if (root.x > root.y || root.y <= 0.001f || root.x >= limit) return result;
    uint32_t stack_node[SG_TERRAIN_STACK];
    float stack_t[SG_TERRAIN_STACK];
    stack_node[0] = 0u;
    stack_t[0] = fmaxf(root.x, 0.0f);
    uint32_t count = 1;
    const uint32_t cells = sc.t.cells;
    while (count != 0u) {
        count -= 1u;
        uint32_t pending = stack_node[count];
        if (stack_t[count] >= limit) continue;
        uint32_t level = first_leading_bit(3u * pending + 1u) / 2u;
        uint32_t size = cells >> level;
        uint32_t morton = pending - ((1u << (2u * level)) - 1u) / 3u;
        uint32_t nx = terrain_compact_bits(morton) * size;
        uint32_t nz = terrain_compact_bits(morton >> 1) * size;
        if (size == 2u) {
            for (uint32_t z = nz; z < nz + 2u; ++z) {
                for (uint32_t x = nx; x < nx + 2u; ++x) {
                    f3 a = terrain_vertex(sc, x, z);
                    f3 b = terrain_vertex(sc, x + 1u, z);
                    f3 c = terrain_vertex(sc, x, z + 1u);
                    f3 e = terrain_vertex(sc, x + 1u, z + 1u);
                    SG_COUNT(SG_EV_TRIANGLE);
                    SG_COUNT(SG_EV_TRIANGLE);
                    f3 first = terrain_triangle(o, d, a, c, b);
                    f3 second = terrain_triangle(o, d, e, b, c);
                    if (first.z < limit) {
                        if (occlusion) return make_float4(0.f, 0.f, 0.f, first.z);
                        f3 n = normalize(terrain_normal(sc, x, z) * (1.0f - first.x - first.y) +
                                         terrain_normal(sc, x, z + 1u) * first.x +
                                         terrain_normal(sc, x + 1u, z) * first.y);
                        result = make_float4(n.x, n.y, n.z, first.z);
                        limit = first.z;
                    }
                    if (second.z < limit) {
                        if (occlusion) return make_float4(0.f, 0.f, 0.f, second.z);
                        f3 n = normalize(terrain_normal(sc, x + 1u, z + 1u) * (1.0f - second.x - second.y) +
                                         terrain_normal(sc, x + 1u, z) * second.x +
                                         terrain_normal(sc, x, z + 1u) * second.y);
                        result = make_float4(n.x, n.y, n.z, second.z);
                        limit = second.z;
                    }
                }
            }
        } else {
            uint32_t children[4];
            float distances[4];
            uint32_t half = size / 2u;
#pragma unroll
            for (uint32_t child = 0; child < 4u; ++child) {
                children[child] = pending * 4u + child + 1u;
                float2 bounds = terrain_interval(sc, o, inv, children[child]);
                SG_COUNT(SG_EV_TERRAIN_BOX);
                distances[child] = (bounds.x <= bounds.y && bounds.y > 0.001f && bounds.x < limit)
                                       ? fmaxf(bounds.x, 0.0f)
                                       : 1.0e30f;
            }
            (void)half;
            // Insertion sort, descending, so the nearest child is pushed last.
#pragma unroll
            for (uint32_t i = 1; i < 4u; ++i) {
#pragma unroll
                for (uint32_t j = i; j > 0u; --j) {
                    if (distances[j - 1u] >= distances[j]) break;
                    float dt = distances[j - 1u];
                    uint32_t ct = children[j - 1u];
                    distances[j - 1u] = distances[j];
                    children[j - 1u] = children[j];
                    distances[j] = dt;
                    children[j] = ct;
                }
            }
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) {
                if (distances[i] < limit && count < SG_TERRAIN_STACK) {
                    stack_node[count] = children[i];
                    stack_t[count] = distances[i];
                    count += 1u;
                }
            }
        }
    }
    return result;
}

// Highest landscape point (src/shaders/ground.wgsl:244-251).
SG_FN float landscape_top(const Scene &sc) {
    float top = sg_u2f(sc.s.globals.ground_counts[3]);
    if (sc.s.globals.ground_counts[1] != 0u) top = fmaxf(top, sc.terrain[sc.t.cache + 5u]);
    return top;
}

// One non-group primitive of the landscape loop (src/shaders/ground.wgsl:179-239).
// Updates `closest`; returns true when an occlusion query is answered.
SG_FN bool prop_test(const Scene &sc, uint32_t i, f3 o, f3 d, Hit &closest, bool occlusion) {
    const sg_primitive &object = sc.objects[i];
    f3 oc = {object.center[0], object.center[1], object.center[2]};
    f3 e = {object.extent[0], object.extent[1], object.extent[2]};
    float kind = object.center[3];
    f3 p = o - oc, v = d;
    f3 angles = f3s(0.f);
    if (kind == 2.25f) {
        angles = rock_angles(object.extent[3]);
        p = rock_to_local(p, angles);
        v = rock_to_local(v, angles);
    } else {
        p = rotate_north(p, -object.extent[3]);
        v = rotate_north(v, -object.extent[3]);
    }
    float distance = 1.0e30f;
    f3 n = f3s(0.f);
    if (kind > 2.5f) {
        float2 bounds = box_interval(p, v, e);
        if (bounds.x > bounds.y || bounds.y <= 0.001f || bounds.x / METRES_PER_AU > closest.distance) return false;
        float4 veg = kind < 3.5f ? conifer_hit(p, v, e) : (kind < 4.5f ? grass_hit(p, v, e) : trunk_hit(p, v, e));
        distance = veg.w;
        n = f3{veg.x, veg.y, veg.z};
    } else if (kind > 1.5f) {
        float2 bounds = box_interval(p, v, e);
        if (bounds.x > bounds.y || bounds.y <= 0.001f || bounds.x / METRES_PER_AU > closest.distance) return false;
        float4 rock = rock_hit(p, v, e);
        distance = rock.w;
        n = f3{rock.x, rock.y, rock.z};
    } else if (kind > 0.5f) {
        f3 sp = p / e, sd = v / e;
        float aa = dot(sd, sd), bb = dot(sp, sd), cc = dot(sp, sp) - 1.0f;
        float hh = bb * bb - aa * cc;
        if (hh >= 0.0f) {
            float root = sqrtf(hh);
            distance = (-bb - root) / aa;
            if (distance <= 0.001f) distance = (-bb + root) / aa;
            n = normalize((p + v * distance) / (e * e));
        }
    } else {
        float2 interval = box_interval(p, v, e);
        if (interval.x <= interval.y && interval.y > 0.001f) {
            distance = interval.x > 0.001f ? interval.x : interval.y;
            f3 point = (p + v * distance) / e;
            f3 ap = fabs3(point);
            if (ap.x >= ap.y && ap.x >= ap.z) n.x = signf(point.x);
            else if (ap.y >= ap.z) n.y = signf(point.y);
            else n.z = signf(point.z);
        }
    }
    if (distance > 0.001f && distance < 1.0e29f && distance / METRES_PER_AU < closest.distance) {
        if (occlusion) {
            closest = Hit{distance / METRES_PER_AU, SG_LOCAL_BASE + 1u + i, f3s(0.f)};
            return true;
        }
        n = kind == 2.25f ? rock_from_local(n, angles) : rotate_north(n, object.extent[3]);
        closest = Hit{distance / METRES_PER_AU, SG_LOCAL_BASE + 1u + i, from_ground(sc, n)};
    }
    return false;
}

// Prop BVH over the leaf primitives (built by the host library from the
// frame's primitive list; group markers only cull, so dropping them keeps
// the closest-hit and any-hit answers of the flat skip-count loop). Nodes:
// lo.xyz + first child / first leaf slot, hi.xyz + leaf count (inner nodes:
// 0x80000000 | split axis; children adjacent). Boxes are tested with the
// loop's own slab test and culling rule; nearer child first. A per-lane
// stack in scratch measured faster on MI355X than stackless (parent-link)
// traversal, which re-tests boxes on the way up.
SG_FN bool prop_bvh_query(const Scene &sc, f3 o, f3 d, Hit &closest, bool occlusion) {
    uint32_t stack[SG_BVH_STACK];
    uint32_t count = 0;
    uint32_t node = 0;
    for (;;) {
        const float4 a = sc.bvh[2 * node], b = sc.bvh[2 * node + 1];
        const uint32_t first = sg_f2u(a.w), leaves = (sg_f2u(b.w) & 0x80000000u) ? 0u : sg_f2u(b.w);
        if (leaves != 0u) {
            for (uint32_t k = 0; k < leaves; ++k) {
                SG_COUNT(SG_EV_PRIM);
                if (prop_test(sc, sc.bvh_prims[first + k], o, d, closest, occlusion)) return true;
            }
        } else {
            float tn[2];
            bool hit[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                SG_COUNT(SG_EV_BVH_BOX);
                const float4 ca = sc.bvh[2 * (first + c)], cb = sc.bvh[2 * (first + c) + 1];
                f3 centre = {0.5f * (ca.x + cb.x), 0.5f * (ca.y + cb.y), 0.5f * (ca.z + cb.z)};
                f3 half = {0.5f * (cb.x - ca.x), 0.5f * (cb.y - ca.y), 0.5f * (cb.z - ca.z)};
                float2 bounds = box_interval(o - centre, d, half);
                hit[c] = !(bounds.x > bounds.y || bounds.y <= 0.001f || bounds.x / METRES_PER_AU > closest.distance);
                tn[c] = bounds.x;
            }
            if (hit[0] && hit[1]) {
                const uint32_t near = tn[1] < tn[0] ? 1u : 0u;
                if (count < SG_BVH_STACK) stack[count++] = first + (near ^ 1u);
                node = first + near;
                continue;
            }
            if (hit[0] || hit[1]) {
                node = first + (hit[0] ? 0u : 1u);
                continue;
            }
        }
        // Pop, skipping nodes the closest hit has since moved in front of.
        bool found = false;
        while (count > 0u) {
            const uint32_t cand = stack[--count];
            SG_COUNT(SG_EV_BVH_BOX);
            const float4 ca = sc.bvh[2 * cand], cb = sc.bvh[2 * cand + 1];
            f3 centre = {0.5f * (ca.x + cb.x), 0.5f * (ca.y + cb.y), 0.5f * (ca.z + cb.z)};
            f3 half = {0.5f * (cb.x - ca.x), 0.5f * (cb.y - ca.y), 0.5f * (cb.z - ca.z)};
            float2 bounds = box_interval(o - centre, d, half);
            if (bounds.x > bounds.y || bounds.y <= 0.001f || bounds.x / METRES_PER_AU > closest.distance) continue;
            node = cand;
            found = true;
            break;
        }
        if (!found) return false;
    }
}

// src/shaders/ground.wgsl:123-242. The host-sphere quadratic runs in FP64.
SG_FN Hit landscape_query(const Scene &sc, const Ray &ray, float limit, bool occlusion) {
    Hit closest{limit, SG_MISS, f3s(0.f)};
    const sg_globals &g = sc.s.globals;
    if (!kLandscape || g.ground_north[3] == 0.0f) return closest;
    SG_COUNT(SG_EV_LANDSCAPE);
    d3 od = ground_origin(sc, ray);
    f3 d = to_ground(sc, ray.direction);
    float radius = g.ground_up[3];
    double rd = radius;
    if (dot(od, od) > 4.0 * rd * rd) return sphere_hit(sc, ray, (uint32_t)g.ground_north[3] - 1u);
    f3 o = tof(od);
    if (d.y >= 0.0f && o.y > landscape_top(sc)) return closest;
    {
        d3 dd = tod(d);
        double a = dot(dd, dd);
        double b = -rd * dd.y - dot(od, dd);
        double c = od.x * od.x + od.z * od.z + od.y * (2.0 * rd + od.y);
        double h = b * b - a * c;
        if (h >= 0.0) {
            double root = sqrt(h);
            double q = b + (b >= 0.0 ? root : -root);
            if (q != 0.0) {
                double t0 = fmin(c / q, q / a), t1 = fmax(c / q, q / a);
                double distance = t0 > 0.001 ? t0 : t1;
                if (distance > 0.001 && (float)(distance / METRES_PER_AU_D) < closest.distance) {
                    if (occlusion) return Hit{(float)(distance / METRES_PER_AU_D), SG_LOCAL_BASE, f3s(0.f)};
                    d3 p = od + dd * distance;
                    d3 n = normalize(d3{p.x / rd, 1.0 + p.y / rd, p.z / rd});
                    closest = Hit{(float)(distance / METRES_PER_AU_D), SG_LOCAL_BASE, from_ground(sc, tof(n))};
                }
            }
        }
    }
    if (g.ground_counts[1] != 0u) {
        float tl = fminf(closest.distance, 1.0e18f) * METRES_PER_AU;
        float4 terrain = terrain_query(sc, o, d, tl, occlusion);
        if (terrain.w < tl && terrain.w < 1.0e29f) {
            closest = Hit{terrain.w / METRES_PER_AU, SG_TERRAIN_ID,
                          from_ground(sc, f3{terrain.x, terrain.y, terrain.z})};
            if (occlusion) return closest;
        }
    }
    if (sc.bvh) {
        if (prop_bvh_query(sc, o, d, closest, occlusion)) return closest;
        return closest;
    }
    for (uint32_t i = 0; i < g.ground_counts[0]; ++i) {
        const sg_primitive &object = sc.objects[i];
        if (object.center[3] < 0.0f) {
            f3 oc = {object.center[0], object.center[1], object.center[2]};
            f3 e = {object.extent[0], object.extent[1], object.extent[2]};
            float2 bounds = box_interval(o - oc, d, e);
            if (bounds.x > bounds.y || bounds.y <= 0.001f || bounds.x / METRES_PER_AU > closest.distance)
                i += (uint32_t)(-object.center[3]);
            continue;
        }
        if (prop_test(sc, i, o, d, closest, occlusion)) return closest;
    }
    return closest;
}

SG_FN Hit landscape_hit(const Scene &sc, const Ray &ray) { return landscape_query(sc, ray, 1.0e30f, false); }

// src/shaders/pathtrace.wgsl:173-185
SG_FN Hit closest_hit(const Scene &sc, const Ray &ray) {
    Hit closest = miss_hit();
    const float host = sc.s.globals.ground_north[3];
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (host == (float)(i + 1u)) continue;
        Hit hit = sphere_hit(sc, ray, i);
        if (hit.distance < closest.distance) closest = hit;
    }
    if (kLandscape && host != 0.0f) {
        Hit local = landscape_hit(sc, ray);
        if (local.distance < closest.distance) closest = local;
    }
    return closest;
}

// --- rings (src/shaders/rings.wgsl) -----------------------------------------
SG_FN RingHit ring_miss(float limit) { return RingHit{limit, SG_MISS, f3s(0.f)}; }
SG_FN RingHit ring_hit(const Scene &sc, const Ray &ray, uint32_t index) {
    SG_COUNT(SG_EV_RING);
    const Body &b = body(sc, index);
    if (b.ring_params.x <= b.ring_plane.w || b.ring_params.y <= 0.0f) return ring_miss(1.0e30f);
    f3 n = xyz(b.ring_plane);
    float denom = dot(ray.direction, n);
    if (fabsf(denom) < 1.0e-7f) return ring_miss(1.0e30f);
    f3 oc = tof(relative_center(sc, ray, index));
    float a = dot(ray.direction, ray.direction);
    f3 perpendicular = cross(ray.direction, cross(ray.direction, oc)) / a;
    float along = -dot(perpendicular, n) / denom;
    float distance = dot(oc, ray.direction) / a + along;
    if (distance <= 0.0f) return ring_miss(1.0e30f);
    f3 point = perpendicular + ray.direction * along;
    float radius = length(point) / b.center.w;
    if (radius < b.ring_plane.w || radius > b.ring_params.x) return ring_miss(1.0e30f);
    return RingHit{distance, index, point};
}
SG_FN RingHit closest_ring(const Scene &sc, const Ray &ray, float limit) {
    RingHit closest = ring_miss(limit);
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        RingHit hit = ring_hit(sc, ray, i);
        if (hit.index != SG_MISS && hit.distance < closest.distance) closest = hit;
    }
    return closest;
}
// Radial bin of the tabulated profile for fractional radius x in [0, 1].
SG_FN uint32_t ring_profile_bin(float x) {
    return umin((uint32_t)fmaxf(x * (float)SG_RING_PROFILE_BINS, 0.0f), SG_RING_PROFILE_BINS - 1u);
}
SG_FN float ring_fraction(const Body &b, f3 point) {
    float radius = length(point) / b.center.w;
    return (radius - b.ring_plane.w) / (b.ring_params.x - b.ring_plane.w);
}
// Analytic banded profile (src/shaders/rings.wgsl:50-81); the optical depth
// comes from the tabulated profile instead for the body that has one.
SG_FN float4 ring_optics(const Scene &sc, uint32_t index, f3 point) {
    const Body &b = body(sc, index);
    float x = ring_fraction(b, point);
    float tau = b.ring_params.y;
    const bool tabulated = index == sc.ring_profile_body;
    f3 color = xyz(b.ring_color);
    if (b.ring_params.z > 0.5f) {
        float broad = 0.5f + 0.5f * sinf(43.0f * x + 1.8f * sinf(17.0f * x));
        float fine = 0.5f + 0.5f * sinf(1900.0f * x + 2.0f * sinf(311.0f * x));
        float micro = 0.5f + 0.5f * sinf(6100.0f * x);
        float density = 0.12f;
        f3 tint = {0.70f, 0.68f, 0.65f};
        if (x >= 0.2792f && x < 0.6910f) {
            density = 1.25f + 1.1f * broad;
            tint = f3s(1.0f);
        } else if (x >= 0.6910f && x < 0.7649f) {
            density = 0.025f;
            tint = f3s(0.75f);
        } else if (x >= 0.7649f) {
            density = 0.48f + 0.22f * broad;
            tint = f3{0.92f, 0.94f, 0.98f};
        }
        density *= 0.65f + 0.45f * fine + 0.12f * micro;
        if (fabsf(x - 0.9487f) < 0.00262f || fabsf(x - 0.99565f) < 0.00034f) density = 0.0f;
        tau *= density;
        color = color * tint * (0.94f + 0.06f * broad);
    }
    if (tabulated) tau = sc.ring_profile[ring_profile_bin(x)];
    return make_float4(color.x, color.y, color.z, tau);
}
SG_FN float ring_transmission(const Scene &sc, const Ray &ray, const RingHit &hit) {
    float mu = fmaxf(fabsf(dot(ray.direction, xyz(body(sc, hit.index).ring_plane))), 1.0e-7f);
    return expf(-ring_optics(sc, hit.index, hit.point).w / mu);
}
SG_FN float one_minus_exp(float x) { return x < 1.0e-3f ? x * (1.0f - 0.5f * x) : 1.0f - expf(-x); }
// Finite-slab single scattering (src/shaders/rings.wgsl:151-166).
SG_FN f3 ring_bsdf(float4 optics, float outgoing, float incoming) {
    float mo = fmaxf(fabsf(outgoing), 1.0e-7f), mi = fmaxf(fabsf(incoming), 1.0e-7f);
    float a = optics.w / mo, b = optics.w / mi;
    f3 c = {optics.x, optics.y, optics.z};
    if (outgoing * incoming < 0.0f) {
        float delta = fabsf(a - b);
        float integral = delta < 1.0e-3f ? 1.0f - 0.5f * delta : one_minus_exp(delta) / fmaxf(delta, 1.0e-20f);
        float factor = optics.w / (mo * mi) * expf(-fminf(a, b)) * integral;
        return c * (factor / (4.0f * PI));
    }
    float factor = one_minus_exp(a + b) / (mo + mi);
    return c * (factor / (4.0f * PI));
}
SG_FN Ray ring_origin(const Scene &sc, const RingHit &hit, f3 direction) {
    const Body &b = body(sc, hit.index);
    f3 n = xyz(b.ring_plane);
    float sgn = dot(direction, n) >= 0.0f ? 1.0f : -1.0f;
    f3 point = hit.point - n * dot(hit.point, n);
    f3 o = point + n * (sgn * b.center.w * 2.0e-6f);
    return Ray{tod(o), direction, hit.index};
}

// Shadow rays (src/shaders/rings.wgsl:90-115).
SG_FN float light_visibility(const Scene &sc, const Ray &ray, uint32_t light) {
    SG_COUNT(SG_EV_SHADOW);
    Hit surface = sphere_hit(sc, ray, light);
    if (surface.index == SG_MISS) return 0.0f;
    const float host = sc.s.globals.ground_north[3];
    if (host == (float)(light + 1u) && closest_hit(sc, ray).index != light) return 0.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (i == light || host == (float)(i + 1u)) continue;
        Hit hit = sphere_hit(sc, ray, i);
        if (hit.index != SG_MISS && (hit.distance < surface.distance || (hit.distance == surface.distance && i < light)))
            return 0.0f;
    }
    if (kLandscape && host != 0.0f) {
        Hit local = landscape_query(sc, ray, surface.distance, true);
        if (local.index != SG_MISS && local.distance < surface.distance) return 0.0f;
    }
    float transmission = 1.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        RingHit hit = ring_hit(sc, ray, i);
        if (hit.index != SG_MISS && hit.distance < surface.distance) transmission *= ring_transmission(sc, ray, hit);
    }
    return transmission;
}
// The shadow ray's crossing of the tabulated ring, if it lies before the
// light: its profile bin and d log(transmission) / d tau (= -1 / mu).
SG_FN bool ring_profile_crossing(const Scene &sc, const Ray &ray, uint32_t light, uint32_t &bin, float &dlog) {
    const uint32_t index = sc.ring_profile_body;
    if (index == SG_MISS) return false;
    RingHit hit = ring_hit(sc, ray, index);
    if (hit.index == SG_MISS) return false;
    Hit surface = sphere_hit(sc, ray, light);
    if (surface.index == SG_MISS || hit.distance >= surface.distance) return false;
    const Body &b = body(sc, index);
    bin = ring_profile_bin(ring_fraction(b, hit.point));
    dlog = -1.0f / fmaxf(fabsf(dot(ray.direction, xyz(b.ring_plane))), 1.0e-7f);
    return true;
}
// ring_bsdf = color * factor / (4 pi): the factor and d factor / d tau.
SG_FN float2 ring_bsdf_factor(float tau, float outgoing, float incoming) {
    float mo = fmaxf(fabsf(outgoing), 1.0e-7f), mi = fmaxf(fabsf(incoming), 1.0e-7f);
    float a = tau / mo, b = tau / mi;
    if (outgoing * incoming < 0.0f) {
        float delta = fabsf(a - b);
        float k = fabsf(1.0f / mo - 1.0f / mi), m = fminf(1.0f / mo, 1.0f / mi);
        float integral, slope;  // I(delta) and dI/d delta
        if (delta < 1.0e-3f) {
            integral = 1.0f - 0.5f * delta;
            slope = -0.5f;
        } else {
            integral = one_minus_exp(delta) / fmaxf(delta, 1.0e-20f);
            slope = (expf(-delta) - integral) / delta;
        }
        float e = expf(-fminf(a, b)) / (mo * mi);
        float factor = tau * e * integral;
        float dfactor = e * (integral - tau * m * integral + tau * k * slope);
        return make_float2(factor, dfactor);
    }
    float q = 1.0f / mo + 1.0f / mi, x = a + b;
    float factor = one_minus_exp(x) / (mo + mi);
    float dfactor = q * (x < 1.0e-3f ? 1.0f - x : expf(-x)) / (mo + mi);
    return make_float2(factor, dfactor);
}
// src/shaders/rings.wgsl:121-141
SG_FN float air_light_visibility(const Scene &sc, const Ray &ray, uint32_t light) {
    const sg_globals &g = sc.s.globals;
    if (!kLandscape || g.ground_north[3] == 0.0f || (float)ground_origin(sc, ray).y <= landscape_top(sc))
        return light_visibility(sc, ray, light);
    SG_COUNT(SG_EV_SHADOW);
    Hit surface = sphere_hit(sc, ray, light);
    if (surface.index == SG_MISS) return 0.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (i == light) continue;
        Hit hit = sphere_hit(sc, ray, i);
        if (hit.index != SG_MISS && (hit.distance < surface.distance || (hit.distance == surface.distance && i < light)))
            return 0.0f;
    }
    float transmission = 1.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        RingHit hit = ring_hit(sc, ray, i);
        if (hit.index != SG_MISS && hit.distance < surface.distance) transmission *= ring_transmission(sc, ray, hit);
    }
    return transmission;
}

// air_light_visibility with its landscape predicates left to the caller,
// for callers that answer them later and coherently (the wavefront
// integrator's k_air_occlusion). Returns the visibility through the bodies
// and rings alone (0 when a body blocks the light) and sets `limit` to the
// light's distance and `owed` to what is left to ask, by which the light is
// blocked: 0 nothing; 1 the landscape before `limit`; light + 2 that too, or
// a closest hit other than the light (the light is the host body). With
// the predicates answered the value is air_light_visibility's: below the
// top of the landscape it runs light_visibility, whose checks are these
// predicates and this loop.
SG_FN float air_light_visibility_split(const Scene &sc, const Ray &ray, uint32_t light, uint32_t &owed,
                                       float &limit) {
    const float host = sc.s.globals.ground_north[3];
    const bool below = kLandscape && host != 0.0f && (float)ground_origin(sc, ray).y <= landscape_top(sc);
    owed = 0u;
    SG_COUNT(SG_EV_SHADOW);
    Hit surface = sphere_hit(sc, ray, light);
    if (surface.index == SG_MISS) return 0.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        if (i == light || (below && host == (float)(i + 1u))) continue;
        Hit hit = sphere_hit(sc, ray, i);
        if (hit.index != SG_MISS && (hit.distance < surface.distance || (hit.distance == surface.distance && i < light)))
            return 0.0f;
    }
    if (below) owed = host == (float)(light + 1u) ? light + 2u : 1u;
    limit = surface.distance;
    float transmission = 1.0f;
    for (uint32_t i = 0; i < body_count(sc); ++i) {
        RingHit hit = ring_hit(sc, ray, i);
        if (hit.index != SG_MISS && hit.distance < surface.distance) transmission *= ring_transmission(sc, ray, hit);
    }
    return transmission;
}
// The predicates air_light_visibility_split left owed: whether they block
// the ray to the light at `limit`.
SG_FN bool air_light_blocked(const Scene &sc, const Ray &ray, float limit, uint32_t owed) {
    if (owed >= 2u && closest_hit(sc, ray).index != owed - 2u) return true;
    Hit local = landscape_query(sc, ray, limit, true);
    return local.index != SG_MISS && local.distance < limit;
}

}  // namespace sg
