// Wavefront integrator kernels, second half: shade, tail, accumulate.
#include "sg_wave.h"

namespace sgk {

__global__ SG_WAVES(SG_SHADE_WAVES) void SG_K(k_shade)(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n,
                                               uint32_t *next, uint32_t *next_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    uint32_t id = 0;
    if (i < n) {
        id = queue[i];
        PathState st;
        load_state(wa, id, st);
        HitRecord h;
        load_hit(wa, id, h);
        // The camera segment was handled by k_camera_air (air on) or has
        // nothing to add (air off).
        st.flags &= ~PATH_CAMERA_SEGMENT;
        alive = path_shade(sc, st, h);
        if (alive) store_state(wa, id, st);
        else finish(wa, id, st);
    }
    // Wave64 compaction: one atomic per wavefront, lane offsets from the ballot.
    const uint64_t mask = __ballot(alive);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(next_count, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (alive) next[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = id;
}

// The tail of a chunk: once fewer paths than the compaction threshold are
// alive, each lane runs its queued path through the remaining loop of
// trace() in one launch (same steps, same draws).
__global__ SG_WAVES(SG_TAIL_WAVES) void SG_K(k_tail)(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    PathState st;
    load_state(wa, id, st);
    run_path(sc, st);
    finish(wa, id, st);
}

// Ordered per-pixel blend of the chunk's path results
// (src/shaders/pathtrace.wgsl:1653-1674; per-sample dispatches when air is on).
__global__ __launch_bounds__(256) void SG_K(k_accumulate)(WaveArgs wa, float4 *accumulation, float4 *backdrop,
                                                    float4 *moments, uint32_t pixels) {
    const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= pixels) return;
    const uint32_t pixel = wa.pixel_begin + local;
    const uint32_t base = local * wa.count;
    float4 mean, bmean;
    float4 m2 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!wa.has_air) {
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f), bsum = sum, sq = sum;
        for (uint32_t k = 0; k < wa.count; ++k) {
            float4 v = wa.value[base + k], b = wa.backdrop[base + k];
            sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
            bsum.x += b.x; bsum.y += b.y; bsum.z += b.z;
            sq.x += v.x * v.x; sq.y += v.y * v.y; sq.z += v.z * v.z;
        }
        const float nf = (float)wa.count;
        mean = make_float4(sum.x / nf, sum.y / nf, sum.z / nf, sum.w / nf);
        bmean = make_float4(bsum.x / nf, bsum.y / nf, bsum.z / nf, 0.f);
        m2 = make_float4(sq.x / nf, sq.y / nf, sq.z / nf, 0.f);
        const uint32_t new_count = wa.accumulated + wa.count;
        if (wa.accumulated > 0u) {
            const float w = nf / (float)new_count;
            float4 old = accumulation[pixel];
            mean = make_float4(old.x + (mean.x - old.x) * w, old.y + (mean.y - old.y) * w,
                               old.z + (mean.z - old.z) * w, old.w + (mean.w - old.w) * w);
            float4 ob = backdrop[pixel];
            bmean = make_float4(ob.x + (bmean.x - ob.x) * w, ob.y + (bmean.y - ob.y) * w, ob.z + (bmean.z - ob.z) * w, 0.f);
            if (moments) {
                float4 om = moments[pixel];
                m2 = make_float4(om.x + (m2.x - om.x) * w, om.y + (m2.y - om.y) * w, om.z + (m2.z - om.z) * w, 0.f);
            }
        }
        m2.w = (float)new_count;
    } else {
        if (wa.accumulated > 0u) {
            mean = accumulation[pixel];
            bmean = backdrop[pixel];
            if (moments) m2 = moments[pixel];
        }
        for (uint32_t k = 0; k < wa.count; ++k) {
            float4 v = wa.value[base + k], b = wa.backdrop[base + k];
            float4 v2 = make_float4(v.x * v.x, v.y * v.y, v.z * v.z, 0.f);
            const uint32_t old_count = wa.accumulated + k;
            if (old_count > 0u) {
                const float w = 1.0f / (float)(old_count + 1u);
                mean = make_float4(mean.x + (v.x - mean.x) * w, mean.y + (v.y - mean.y) * w,
                                   mean.z + (v.z - mean.z) * w, mean.w + (v.w - mean.w) * w);
                bmean = make_float4(bmean.x + (b.x - bmean.x) * w, bmean.y + (b.y - bmean.y) * w,
                                    bmean.z + (b.z - bmean.z) * w, 0.f);
                m2 = make_float4(m2.x + (v2.x - m2.x) * w, m2.y + (v2.y - m2.y) * w, m2.z + (v2.z - m2.z) * w, 0.f);
            } else {
                mean = v;
                bmean = make_float4(b.x, b.y, b.z, 0.f);
                m2 = v2;
            }
        }
        bmean.w = 0.f;
        m2.w = (float)(wa.accumulated + wa.count);
    }
    accumulation[pixel] = mean;
    backdrop[pixel] = bmean;
    if (moments) moments[pixel] = m2;
}

}  // namespace sgk
