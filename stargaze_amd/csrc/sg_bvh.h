// Host-side builder of the prop BVH traversed by prop_bvh_query.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/stargaze_amd.h"
#include "sg_math.h"

namespace sg_host {

// --- prop BVH ---------------------------------------------------------------
// Binary BVH over the non-group primitives of the frame's list. Leaf boxes
// are the world box of each primitive's local extent box under its rotation
// (north-axis angle, or the packed Rz*Ry*Rx of tumbled rocks; the same
// conservative box src/ground.rs:172-196 uses for its groups) plus the
// reference's 0.02 m group margin. Every primitive kind lies inside its
// extent box or is culled by it in the shader, so the BVH changes no answer.
constexpr uint32_t kBvhLeaf = 4;

inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
constexpr size_t kBvhMinPrims = 16;

struct BvhPrim {
    float lo[3], hi[3], c[3];
    uint32_t index;
};

inline void prim_bounds(const sg_primitive &p, BvhPrim &out) {
    double r[3][3];
    if (p.center[3] == 2.25f) {
        uint32_t bits = (uint32_t)p.extent[3];
        const double k = 6.283185307179586 / 256.0;
        double ax = (bits & 255u) * k, ay = ((bits >> 8) & 255u) * k, az = ((bits >> 16) & 255u) * k;
        double cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay), cz = std::cos(az),
               sz = std::sin(az);
        // Rz(az) * Ry(ay) * Rx(ax)
        double rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
        double ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
        double rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
        double t[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                t[i][j] = 0;
                for (int m = 0; m < 3; ++m) t[i][j] += ry[i][m] * rx[m][j];
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                r[i][j] = 0;
                for (int m = 0; m < 3; ++m) r[i][j] += rz[i][m] * t[m][j];
            }
    } else {
        double c = std::cos((double)p.extent[3]), s = std::sin((double)p.extent[3]);
        double rz[3][3] = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
        std::memcpy(r, rz, sizeof(r));
    }
    for (int i = 0; i < 3; ++i) {
        double h = 0;
        for (int j = 0; j < 3; ++j) h += std::fabs(r[i][j]) * p.extent[j];
        h = h * (1.0 + 1e-5) + 0.02;
        out.lo[i] = (float)(p.center[i] - h);
        out.hi[i] = (float)(p.center[i] + h);
        out.c[i] = p.center[i];
    }
}

// Median split on the longest centroid axis; children stored adjacently.
inline void build_bvh(const sg_primitive *objects, uint32_t n, std::vector<float4> &nodes,
                      std::vector<uint32_t> &order) {
    std::vector<BvhPrim> prims;
    prims.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (objects[i].center[3] < 0.0f) continue;
        BvhPrim b;
        prim_bounds(objects[i], b);
        b.index = i;
        prims.push_back(b);
    }
    nodes.clear();
    order.clear();
    nodes.resize(2);
    if (prims.empty()) {
        nodes[0] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, u2f(0u));
        nodes[1] = make_float4(-1.0e30f, -1.0e30f, -1.0e30f, u2f(0u));
        return;
    }
    struct Task { uint32_t node, begin, end; };
    std::vector<Task> tasks{{0u, 0u, (uint32_t)prims.size()}};
    while (!tasks.empty()) {
        Task t = tasks.back();
        tasks.pop_back();
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
        for (uint32_t k = t.begin; k < t.end; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], prims[k].lo[a]);
                hi[a] = std::max(hi[a], prims[k].hi[a]);
                clo[a] = std::min(clo[a], prims[k].c[a]);
                chi[a] = std::max(chi[a], prims[k].c[a]);
            }
        const uint32_t count = t.end - t.begin;
        if (count <= kBvhLeaf) {
            nodes[2 * t.node] = make_float4(lo[0], lo[1], lo[2], u2f((uint32_t)order.size()));
            nodes[2 * t.node + 1] = make_float4(hi[0], hi[1], hi[2], u2f(count));
            for (uint32_t k = t.begin; k < t.end; ++k) order.push_back(prims[k].index);
            continue;
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        const uint32_t mid = t.begin + count / 2;
        std::nth_element(prims.begin() + t.begin, prims.begin() + mid, prims.begin() + t.end,
                         [axis](const BvhPrim &x, const BvhPrim &y) { return x.c[axis] < y.c[axis]; });
        const uint32_t left = (uint32_t)(nodes.size() / 2);
        nodes.resize(nodes.size() + 4);
        nodes[2 * t.node] = make_float4(lo[0], lo[1], lo[2], u2f(left));
        nodes[2 * t.node + 1] = make_float4(hi[0], hi[1], hi[2], u2f(0x80000000u | (uint32_t)axis));
        tasks.push_back({left + 1, mid, t.end});
        tasks.push_back({left, t.begin, mid});
    }
}

}  // namespace sg_host
