// Device blocks shared by the fused, batched and adjoint kernels: one block
// of pixel samples (trace_block) and the per-pixel blend (blend_pixel).
#pragma once
#include "sg_launch.h"

namespace sgk {

// The per-pixel epilogue of `main` (src/shaders/pathtrace.wgsl:1653-1674):
// ordered mean of this pass's samples and the running-average blend, or,
// with air, one blend per sample (src/pathtracer.rs:861-882,971-976).
static __device__ __forceinline__ void blend_pixel(const TraceArgs &ta, uint32_t pixel, const float4 *values,
                                            const float4 *backdrops) {
    const uint32_t base = 0;

    float4 mean, bmean;
    float4 m2 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!ta.has_air) {
        // One dispatch of counts.z samples (src/shaders/pathtrace.wgsl:1653-1674).
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f), bsum = sum, sq = sum;
        for (uint32_t k = 0; k < ta.count; ++k) {
            float4 v = values[base + k], b = backdrops[base + k];
            sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
            bsum.x += b.x; bsum.y += b.y; bsum.z += b.z;
            sq.x += v.x * v.x; sq.y += v.y * v.y; sq.z += v.z * v.z;
        }
        const float n = (float)ta.count;
        mean = make_float4(sum.x / n, sum.y / n, sum.z / n, sum.w / n);
        bmean = make_float4(bsum.x / n, bsum.y / n, bsum.z / n, 0.f);
        m2 = make_float4(sq.x / n, sq.y / n, sq.z / n, 0.f);
        const uint32_t new_count = ta.accumulated + ta.count;
        if (ta.accumulated > 0u) {
            const float w = n / (float)new_count;
            float4 old = ta.accumulation[pixel];
            mean = make_float4(old.x + (mean.x - old.x) * w, old.y + (mean.y - old.y) * w,
                               old.z + (mean.z - old.z) * w, old.w + (mean.w - old.w) * w);
            float4 ob = ta.backdrop[pixel];
            bmean = make_float4(ob.x + (bmean.x - ob.x) * w, ob.y + (bmean.y - ob.y) * w,
                                ob.z + (bmean.z - ob.z) * w, 0.f);
            if (ta.moments) {
                float4 om = ta.moments[pixel];
                m2 = make_float4(om.x + (m2.x - om.x) * w, om.y + (m2.y - om.y) * w, om.z + (m2.z - om.z) * w, 0.f);
            }
        }
        m2.w = (float)new_count;
    } else {
        // One dispatch per sample with counts.z = 1 (src/pathtracer.rs:861-882,971-976).
        if (ta.accumulated > 0u) {
            mean = ta.accumulation[pixel];
            bmean = ta.backdrop[pixel];
            if (ta.moments) m2 = ta.moments[pixel];
        }
        for (uint32_t k = 0; k < ta.count; ++k) {
            float4 v = values[base + k], b = backdrops[base + k];
            float4 v2 = make_float4(v.x * v.x, v.y * v.y, v.z * v.z, 0.f);
            const uint32_t old_count = ta.accumulated + k;
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
        m2.w = (float)(ta.accumulated + ta.count);
    }
    ta.accumulation[pixel] = mean;
    ta.backdrop[pixel] = bmean;
    if (ta.moments) ta.moments[pixel] = m2;
}
// One block of pixel samples: lane t traces sample t % count of pixel
// block * pixels_per_block + t / count, then one lane per pixel forms the
// ordered mean and blend. Every lane of the workgroup must call it.
static __device__ __forceinline__ void trace_block(const Scene &sc, const TraceArgs &ta, uint32_t block, float4 *s_value,
                                            float4 *s_backdrop) {
    const uint32_t t = threadIdx.x;
    const uint32_t lp = t / ta.count;
    const uint32_t s = t - lp * ta.count;
    const uint32_t pixel = block * ta.pixels_per_block + lp;
    const uint32_t pixels = ta.width * ta.height;
    const bool active = lp < ta.pixels_per_block && pixel < pixels;
    float4 value = make_float4(0.f, 0.f, 0.f, 0.f);
    f3 backdrop = f3s(0.f);
    if (active) {
        const uint32_t x = pixel % ta.width, y = pixel / ta.width;
        const uint32_t accumulated = ta.has_air ? ta.accumulated + s : ta.accumulated;
        uint32_t rng = path_seed(pixel, ta.sequence + s);
        f3 direction = camera_direction(sc, x, y, rng);
        PathState st;
        path_begin(st, direction, rng, accumulated);
        run_path(sc, st);
        PathResult r = path_finish(st);
        value = r.value;
        backdrop = r.backdrop;
        if (ta.records)
            ta.records[(size_t)pixel * ta.count + s] = make_float4(value.x, value.y, value.z, ring_cap_scale(st));
    }
    s_value[t] = value;
    s_backdrop[t] = make_float4(backdrop.x, backdrop.y, backdrop.z, 0.f);
    __syncthreads();
    if (!active || s != 0u) return;
    blend_pixel(ta, pixel, s_value + lp * ta.count, s_backdrop + lp * ta.count);
}

// One pass of a batch (see BatchArgs): the persistent loop over pixel blocks.
template <uint32_t B>
// The fused kernel's occupancy budget: without it the persistent loop's few
// extra registers would halve the occupancy of the frames without a ground
// scene.
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(SG_FUSED_WAVES, 8))) void SG_K(k_batch)(
    BatchArgs ba) {
    __shared__ float4 s_value[B];
    __shared__ float4 s_backdrop[B];
    __shared__ uint32_t s_begin[kMaxBatchViews + 1];
    __shared__ uint32_t s_first;
    // The table does not change while the kernel runs. Read through the
    // constant address space, a view's scene and pass constants come in by
    // scalar loads, as kernel arguments do for k_trace; through a plain
    // global pointer every load after the first stores of the loop would be
    // a vector load of a uniform value.
    using ConstView = __attribute__((address_space(4))) const BatchView;
    const ConstView *table = (const ConstView *)ba.table;
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < ba.entries; e += B) s_begin[e] = table[e].block_begin;
    if (t == 0) {
        s_begin[ba.entries] = ba.total_blocks;
        atomicMin(&ba.clock[0], (unsigned long long)wall_clock64());
    }
    for (;;) {
        __syncthreads();
        if (t == 0) s_first = atomicAdd(&ba.counter[0], ba.blocks_per_pull);
        __syncthreads();
        const uint32_t first = s_first;
        if (first >= ba.total_blocks) break;
        const uint32_t last = umin(first + ba.blocks_per_pull, ba.total_blocks);
        // Entry owning the first block: last e with s_begin[e] <= first. The
        // blocks of a pull are consecutive, so later ones step to the next
        // entry instead of searching again.
        uint32_t lo = 0, hi = ba.entries;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_begin[mid] <= first) lo = mid;
            else hi = mid;
        }
        uint32_t e = __builtin_amdgcn_readfirstlane(lo);
        for (uint32_t block = first; block < last; ++block) {
            while (block >= s_begin[e + 1u]) {
                // The entry's finish time is its last block's: a max that
                // returns nothing, so no wave waits on it.
                if (t == 0) atomicMax(&ba.done[e], (unsigned long long)wall_clock64());
                e = __builtin_amdgcn_readfirstlane(e + 1u);
            }
            const BatchView &v = *(const BatchView *)&table[e];
            trace_block(v.scene, v.ta, block - s_begin[e], s_value, s_backdrop);
            __syncthreads();
        }
        if (t == 0) atomicMax(&ba.done[e], (unsigned long long)wall_clock64());
    }
    if (t == 0) atomicMax(&ba.clock[1], (unsigned long long)wall_clock64());
}

}  // namespace sgk
