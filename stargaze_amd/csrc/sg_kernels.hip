// HIP kernels and the C ABI of the MI355X path tracer.
//
// Launch structure (per progressive pass, one HIP stream per context):
//   k_terrain_cache   once per terrain upload / host radius change
//                     (cache_terrain_geometry, src/shaders/terrain.wgsl:242-285)
//   k_sky_cache       on a fresh accumulation (cache_surface_sky, forest.wgsl:3-17)
//   k_trace           every pass: W*H*count path samples, one lane per
//                     (pixel, sample); lanes of one pixel share a wavefront,
//                     and the per-pixel mean and running-average blend of
//                     src/shaders/pathtrace.wgsl:1645-1675 are formed in LDS in
//                     the reference's summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sg_bvh.h"
#include "sg_integrator.h"

using namespace sg;

namespace {

constexpr uint32_t kTraceBlock = 256;

// Register budget of the heavy wavefront kernels, as waves per SIMD.
#ifndef SG_HEAVY_WAVES
#define SG_HEAVY_WAVES 1
#endif
#define SG_HEAVY __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SG_HEAVY_WAVES, 8)))

struct TraceArgs {
    float4 *accumulation;
    float4 *backdrop;
    float4 *moments;       // null unless second moments are kept
    float4 *records;       // per path sample (pixel * count + s): radiance rgb, ring-cap scale; or null
    uint32_t width, height;
    uint32_t count;        // samples this pass (counts.z)
    uint32_t accumulated;  // samples already in the image (counts.y)
    uint32_t sequence;     // noise sequence (sampling.x)
    uint32_t has_air;      // per-sample dispatch semantics
    uint32_t pixels_per_block;
};

// The per-pixel epilogue of `main` (src/shaders/pathtrace.wgsl:1653-1674):
// ordered mean of this pass's samples and the running-average blend, or,
// with air, one blend per sample (src/pathtracer.rs:861-882,971-976).
__device__ __forceinline__ void blend_pixel(const TraceArgs &ta, uint32_t pixel, const float4 *values,
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
__device__ __forceinline__ void trace_block(const Scene &sc, const TraceArgs &ta, uint32_t block, float4 *s_value,
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

__global__ __launch_bounds__(kTraceBlock) void k_trace(Scene sc, TraceArgs ta) {
    __shared__ float4 s_value[kTraceBlock];
    __shared__ float4 s_backdrop[kTraceBlock];
    trace_block(sc, ta, blockIdx.x, s_value, s_backdrop);
}


// ---------------------------------------------------------------------------
// Wavefront integrator: the loop of trace() run as one kernel per step over
// a queue of live paths (structure-of-arrays state, wave64 ballot
// compaction), then an ordered per-pixel accumulation.
// ---------------------------------------------------------------------------
struct WaveArgs {
    // Path state, indexed by path id within the chunk.
    double4 *ray_o;    // offset.xyz, anchor bits in .w
    float4 *ray_d;     // direction.xyz, previous_pdf
    double4 *prev_o;   // previous vertex offset.xyz, anchor bits
    float4 *thr;       // throughput.xyz, ring_cap
    float4 *rad;       // radiance.xyz, shine_body bits
    float4 *rbase;     // ring_base.xyz, flags | depth << 8 | null_crossings << 16
    uint32_t *rng;
    // Hit records from path_extend.
    float4 *h0;        // hit.distance, hit.index bits, disc_scale, ring.distance
    float4 *h1;        // hit.normal.xyz, ring.index bits
    float4 *h2;        // ring.point.xyz
    // Per-path results.
    float4 *value;
    float4 *backdrop;
    float4 *records;   // per path sample of the pass (see TraceArgs), or null
    uint32_t width, height;
    uint32_t count, accumulated, sequence, has_air;
    uint32_t pixel_begin, paths;
};

SG_FN double bits_d(uint32_t u) { return (double)u; }
SG_FN uint32_t d_bits(double d) { return (uint32_t)d; }

__device__ __forceinline__ uint32_t path_accumulated(const WaveArgs &wa, uint32_t id) {
    return wa.has_air ? wa.accumulated + id % wa.count : wa.accumulated;
}
__device__ __forceinline__ void load_ray(const double4 &o, const float4 &d, Ray &r) {
    r.offset = d3{o.x, o.y, o.z};
    r.anchor = d_bits(o.w);
    r.direction = f3{d.x, d.y, d.z};
}
__device__ __forceinline__ void load_state(const WaveArgs &wa, uint32_t id, PathState &st) {
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
__device__ __forceinline__ void store_state(const WaveArgs &wa, uint32_t id, const PathState &st) {
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
__device__ __forceinline__ void load_hit(const WaveArgs &wa, uint32_t id, HitRecord &h) {
    float4 a = wa.h0[id], b = wa.h1[id], c = wa.h2[id];
    h.hit.distance = a.x;
    h.hit.index = __float_as_uint(a.y);
    h.disc_scale = a.z;
    h.ring.distance = a.w;
    h.hit.normal = f3{b.x, b.y, b.z};
    h.ring.index = __float_as_uint(b.w);
    h.ring.point = f3{c.x, c.y, c.z};
}
__device__ __forceinline__ void finish(const WaveArgs &wa, uint32_t id, const PathState &st) {
    PathResult r = path_finish(st);
    wa.value[id] = r.value;
    wa.backdrop[id] = make_float4(r.backdrop.x, r.backdrop.y, r.backdrop.z, 0.f);
    if (wa.records)
        wa.records[(size_t)wa.pixel_begin * wa.count + id] =
            make_float4(r.value.x, r.value.y, r.value.z, ring_cap_scale(st));
}

__global__ __launch_bounds__(256) void k_generate(Scene sc, WaveArgs wa, uint32_t *queue) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= wa.paths) return;
    const uint32_t pixel = wa.pixel_begin + id / wa.count;
    const uint32_t s = id % wa.count;
    const uint32_t x = pixel % wa.width, y = pixel / wa.width;
    uint32_t rng = path_seed(pixel, wa.sequence + s);
    f3 direction = camera_direction(sc, x, y, rng);
    PathState st;
    path_begin(st, direction, rng, path_accumulated(wa, id));
    store_state(wa, id, st);
    queue[id] = id;
}

__global__ SG_HEAVY void k_extend(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    PathState st;
    load_ray(wa.ray_o[id], wa.ray_d[id], st.ray);
    uint32_t packed = __float_as_uint(wa.rbase[id].w);
    st.depth = (packed >> 8) & 255u;
    st.null_crossings = packed >> 16;
    st.rng = wa.rng[id];
    HitRecord h = path_extend(sc, st);
    wa.rng[id] = st.rng;
    wa.h0[id] = make_float4(h.hit.distance, __uint_as_float(h.hit.index), h.disc_scale, h.ring.distance);
    wa.h1[id] = make_float4(h.hit.normal.x, h.hit.normal.y, h.hit.normal.z, __uint_as_float(h.ring.index));
    wa.h2[id] = make_float4(h.ring.point.x, h.ring.point.y, h.ring.point.z, 0.f);
}

__global__ SG_HEAVY void k_camera_air(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = queue[i];
    PathState st;
    float4 d = wa.ray_d[id];
    load_ray(wa.ray_o[id], d, st.ray);
    float4 t = wa.thr[id], r = wa.rad[id], b = wa.rbase[id];
    st.throughput = f3{t.x, t.y, t.z};
    st.radiance = f3{r.x, r.y, r.z};
    st.flags = __float_as_uint(b.w) & 255u;
    st.rng = wa.rng[id];
    st.accumulated = path_accumulated(wa, id);
    HitRecord h;
    load_hit(wa, id, h);
    path_camera_air(sc, st, h);
    wa.thr[id] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, t.w);
    wa.rad[id] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, r.w);
    wa.rbase[id] = make_float4(b.x, b.y, b.z, __uint_as_float((__float_as_uint(b.w) & ~255u) | st.flags));
    wa.rng[id] = st.rng;
}

__global__ SG_HEAVY void k_shade(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n,
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
__global__ SG_HEAVY void k_tail(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n) {
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
__global__ __launch_bounds__(256) void k_accumulate(WaveArgs wa, float4 *accumulation, float4 *backdrop,
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

// ---------------------------------------------------------------------------
// Batched views: one persistent launch renders a progressive pass of many
// views (own scene, resolution, sample count). The views' pixel blocks are
// laid end to end in table order; workgroups pull runs of blocks from an
// atomic counter and trace each with the fused integrator's trace_block,
// reading the view's scene from device memory.
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxBatchViews = 1024;

struct BatchView {
    Scene scene;
    TraceArgs ta;          // pixels_per_block = workgroup size / samples
    uint32_t block_begin;  // first block of this view in the pass
    uint32_t blocks;
    uint32_t view;         // index of the view in the batch
    uint32_t pad;
};

struct BatchArgs {
    BatchView *table;             // views with samples to take, in launch order
    uint32_t entries;
    uint32_t total_blocks;
    uint32_t blocks_per_pull;
    uint32_t *counter;            // [0] next block; [1 + e] blocks left of entry e
    unsigned long long *done;     // [e] clock at which entry e finished
    unsigned long long *clock;    // [0] first workgroup start, [1] last workgroup end
};

__global__ void k_batch_reset(BatchArgs ba) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ba.counter[0] = 0u;
        ba.clock[0] = ~0ull;
        ba.clock[1] = 0ull;
    }
    if (i < ba.entries) {
        ba.counter[1 + i] = ba.table[i].blocks;
        ba.done[i] = 0ull;
    }
}

template <uint32_t B>
__global__ __launch_bounds__(B) void k_batch(BatchArgs ba) {
    __shared__ float4 s_value[B];
    __shared__ float4 s_backdrop[B];
    __shared__ uint32_t s_begin[kMaxBatchViews + 1];
    __shared__ uint32_t s_first;
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < ba.entries; e += B) s_begin[e] = ba.table[e].block_begin;
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
        for (uint32_t block = first; block < last; ++block) {
            // Entry owning the block: last e with s_begin[e] <= block.
            uint32_t lo = 0, hi = ba.entries;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_begin[mid] <= block) lo = mid;
                else hi = mid;
            }
            const uint32_t e = __builtin_amdgcn_readfirstlane(lo);
            const BatchView &v = ba.table[e];
            trace_block(v.scene, v.ta, block - s_begin[e], s_value, s_backdrop);
            __syncthreads();
            if (t == 0) {
                __threadfence();
                if (atomicSub(&ba.counter[1 + e], 1u) == 1u) ba.done[e] = wall_clock64();
            }
        }
    }
    if (t == 0) atomicMax(&ba.clock[1], (unsigned long long)wall_clock64());
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

// ---------------------------------------------------------------------------
// Adjoint of one progressive pass (no reference counterpart). Lane layout as
// k_trace; each lane runs its path sample's primal and replay (adjoint_path),
// adding d(loss)/d(parameters) into a workgroup copy of the gradient that is
// flushed to the device gradient once, at the end of the persistent loop.
// ---------------------------------------------------------------------------
struct AdjointArgs {
    TraceArgs ta;            // pass layout; accumulation/backdrop receive the pass mean (or null)
    const float4 *recorded;  // the pass's recorded primal outputs (skip the primal runs), or null
    const float *d_image;    // d(loss)/d(pass mean), rgb per pixel
    float *gradient;         // GRAD_COUNT
    uint32_t blocks;
};

__global__ void k_fill(float *p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kTraceBlock) void k_adjoint(Scene sc, AdjointArgs aa) {
    __shared__ float s_grad[GRAD_COUNT];
    __shared__ float s_density[GRAD_PRIVATE * kTraceBlock];  // lane-private, slot-major
    __shared__ float4 s_value[kTraceBlock];
    __shared__ float4 s_backdrop[kTraceBlock];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < GRAD_COUNT; i += blockDim.x) s_grad[i] = 0.0f;
    for (uint32_t b = 0; b < GRAD_PRIVATE; ++b) s_density[b * blockDim.x + t] = 0.0f;
    __syncthreads();
    const TraceArgs &ta = aa.ta;
    const uint32_t lp = t / ta.count;
    const uint32_t s = t - lp * ta.count;
    const uint32_t pixels = ta.width * ta.height;
    const float inv_count = 1.0f / (float)ta.count;
    for (uint32_t block = blockIdx.x; block < aa.blocks; block += gridDim.x) {
        const uint32_t pixel = block * ta.pixels_per_block + lp;
        const bool active = lp < ta.pixels_per_block && pixel < pixels;
        float4 value = make_float4(0.f, 0.f, 0.f, 0.f);
        f3 backdrop = f3s(0.f);
        if (active) {
            const uint32_t x = pixel % ta.width, y = pixel / ta.width;
            const uint32_t accumulated = ta.has_air ? ta.accumulated + s : ta.accumulated;
            uint32_t rng = path_seed(pixel, ta.sequence + s);
            f3 direction = camera_direction(sc, x, y, rng);
            const float *d = aa.d_image + 3u * (size_t)pixel;
            f3 w = f3{d[0], d[1], d[2]} * inv_count;
            if (aa.recorded) {
                const float4 rec = aa.recorded[(size_t)pixel * ta.count + s];
                replay_path(sc, direction, rng, accumulated, w, f3{rec.x, rec.y, rec.z}, rec.w, s_grad,
                            s_density + t, blockDim.x);
            } else {
                PathResult r = adjoint_path(sc, direction, rng, accumulated, w, s_grad, s_density + t, blockDim.x);
                value = r.value;
                backdrop = r.backdrop;
            }
        }
        if (ta.accumulation) {
            s_value[t] = value;
            s_backdrop[t] = make_float4(backdrop.x, backdrop.y, backdrop.z, 0.f);
            __syncthreads();
            if (active && s == 0u) {
                TraceArgs mean = ta;
                mean.accumulated = 0u;
                mean.moments = nullptr;
                blend_pixel(mean, pixel, s_value + lp * ta.count, s_backdrop + lp * ta.count);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < GRAD_COUNT; i += blockDim.x)
        if (s_grad[i] != 0.0f) atomicAdd(&aa.gradient[i], s_grad[i]);
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
    case SG_PROBE_CONE: {
        r.distance = cone_width(sc, ray, pr.anchor == SG_MISS ? 0u : pr.anchor);
        break;
    }
    default:
        break;
    }
    out[i] = r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host library: the PathTracer state machine of src/pathtracer.rs:164-1009.
// ---------------------------------------------------------------------------
struct sg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t width = 0, height = 0;
    sg_options options{};
    // History (src/pathtracer.rs:164-188)
    std::vector<uint8_t> key;
    uint32_t samples = 0, sequence = 0;
    uint32_t exposure_idle_frames = 0;
    bool exposure_was_enabled = false;
    // Device buffers
    float4 *accumulation = nullptr, *backdrop = nullptr, *moments = nullptr;
    bool moments_on = false;
    Body *bodies = nullptr;
    sg_primitive *objects = nullptr;
    uint32_t object_capacity = 0;
    float4 *sky_light = nullptr;
    float *terrain = nullptr;
    uint64_t terrain_floats = 0;
    uint32_t terrain_cells = 0;
    bool terrain_cached = false;
    uint32_t terrain_radius_bits = 0;
    sg_ray_star *catalogue = nullptr;
    uint32_t catalogue_count = 0;
    uint32_t *sky_cells = nullptr;
    uint32_t *milky_way = nullptr;
    uint32_t milky_w = 0, milky_h = 0;
    sg_probe_ray *probe_rays = nullptr;
    sg_probe_result *probe_results = nullptr;
    uint32_t probe_capacity = 0;
    // Prop BVH, rebuilt when the primitive list changes.
    std::vector<uint8_t> bvh_key;
    float4 *bvh_nodes = nullptr;
    uint32_t *bvh_prims = nullptr;
    size_t bvh_node_capacity = 0, bvh_prim_capacity = 0;
    bool bvh_valid = false;
    // Integrator selection and wavefront buffers
    uint32_t integrator = SG_INTEGRATOR_WAVEFRONT;
    uint32_t chunk_paths = 1u << 23;
    uint32_t tail_paths = 0;  // compaction threshold: fused tail below this many live paths
    uint32_t wave_capacity = 0;
    void *wave_block = nullptr;
    WaveArgs wave{};
    uint32_t *queues[2] = {nullptr, nullptr};
    uint32_t *counter = nullptr;
    uint32_t *counter_host = nullptr;
    // Measurement
    bool profiling = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    double kernel_ms = 0.0;
    uint64_t launches = 0;
    std::string error;
    // Device copies of the last encoded frame's bodies and primitives.
    bool frame_on_device = false;
    // Batch this context is a view of, and the event of its last queued pass,
    // which work on the context stream must wait for.
    sg_batch *batch = nullptr;
    hipEvent_t batch_event = nullptr;
    hipEvent_t ev_ready = nullptr;
    // Tabulated ring profile (SG_RING_PROFILE_BINS floats) and its body.
    float *ring_profile = nullptr;
    uint32_t ring_profile_body = SG_MISS;
    // Tabulated air density (SG_DENSITY_BINS floats) while density_on.
    float *density_profile = nullptr;
    bool density_on = false;
    // The last queued pass, which sg_adjoint differentiates.
    struct {
        uint32_t count = 0, samples = 0, sequence = 0;
        bool recorded = false;  // its primal outputs are in `records`
    } last;
    // Per path sample primal outputs of the last pass, kept while record_on.
    bool record_on = false;
    float4 *records = nullptr;
    uint64_t record_capacity = 0;
    // Adjoint buffers: d(loss)/d(image) rgb, gradient, recomputed pass mean.
    float *adj_image = nullptr;
    float *adj_gradient = nullptr;
    float4 *adj_primal = nullptr;
    uint64_t adj_pixels = 0;
    double adjoint_ms = 0.0;
    uint64_t adjoint_launches = 0;
};

namespace {

int fail(sg_ctx *ctx, int code, const std::string &message) {
    if (ctx) ctx->error = message;
    return code;
}

// Order the context stream after the last batch pass that rendered this view.
void join_batch(sg_ctx *ctx) {
    if (ctx->batch_event) {
        (void)hipStreamWaitEvent(ctx->stream, ctx->batch_event, 0);
        ctx->batch_event = nullptr;
    }
}
#define SG_HIP(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), SG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

int alloc_image(sg_ctx *ctx) {
    const size_t bytes = std::max<size_t>((size_t)ctx->width * ctx->height * 16u, 16u);
    if (ctx->accumulation) (void)hipFree(ctx->accumulation);
    if (ctx->backdrop) (void)hipFree(ctx->backdrop);
    if (ctx->moments) (void)hipFree(ctx->moments);
    ctx->accumulation = ctx->backdrop = ctx->moments = nullptr;
    SG_HIP(ctx, hipMalloc(&ctx->accumulation, bytes));
    SG_HIP(ctx, hipMalloc(&ctx->backdrop, bytes));
    SG_HIP(ctx, hipMemsetAsync(ctx->accumulation, 0, bytes, ctx->stream));
    SG_HIP(ctx, hipMemsetAsync(ctx->backdrop, 0, bytes, ctx->stream));
    if (ctx->moments_on) {
        SG_HIP(ctx, hipMalloc(&ctx->moments, bytes));
        SG_HIP(ctx, hipMemsetAsync(ctx->moments, 0, bytes, ctx->stream));
    }
    return SG_OK;
}

// Rebuild and upload the BVH when the primitive list changed.
int ensure_bvh(sg_ctx *ctx, const sg_frame *frame) {
    if (frame->ground_count < sg_host::kBvhMinPrims) {
        ctx->bvh_valid = false;
        return SG_OK;
    }
    const uint8_t *bytes = reinterpret_cast<const uint8_t *>(frame->ground);
    const size_t n = (size_t)frame->ground_count * sizeof(sg_primitive);
    if (ctx->bvh_valid && ctx->bvh_key.size() == n && std::memcmp(ctx->bvh_key.data(), bytes, n) == 0) return SG_OK;
    std::vector<float4> nodes;
    std::vector<uint32_t> order;
    sg_host::build_bvh(frame->ground, frame->ground_count, nodes, order);
    if (order.empty()) order.push_back(0u);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (nodes.size() > ctx->bvh_node_capacity) {
        if (ctx->bvh_nodes) SG_HIP(ctx, hipFree(ctx->bvh_nodes));
        ctx->bvh_nodes = nullptr;
        SG_HIP(ctx, hipMalloc(&ctx->bvh_nodes, nodes.size() * sizeof(float4)));
        ctx->bvh_node_capacity = nodes.size();
    }
    if (order.size() > ctx->bvh_prim_capacity) {
        if (ctx->bvh_prims) SG_HIP(ctx, hipFree(ctx->bvh_prims));
        ctx->bvh_prims = nullptr;
        SG_HIP(ctx, hipMalloc(&ctx->bvh_prims, order.size() * sizeof(uint32_t)));
        ctx->bvh_prim_capacity = order.size();
    }
    SG_HIP(ctx, hipMemcpy(ctx->bvh_nodes, nodes.data(), nodes.size() * sizeof(float4), hipMemcpyHostToDevice));
    SG_HIP(ctx, hipMemcpy(ctx->bvh_prims, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    ctx->bvh_key.assign(bytes, bytes + n);
    ctx->bvh_valid = true;
    return SG_OK;
}

Scene make_scene(const sg_ctx *ctx, const sg_settings &settings) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = settings;
    sc.bodies = ctx->bodies;
    sc.objects = ctx->objects;
    sc.sky_light = ctx->sky_light;
    sc.terrain = ctx->terrain;
    sc.catalogue = ctx->catalogue;
    sc.sky_cells = ctx->sky_cells;
    sc.catalogue_count = ctx->catalogue_count;
    sc.milky_way = ctx->milky_way;
    sc.bvh = ctx->bvh_valid ? ctx->bvh_nodes : nullptr;
    sc.bvh_prims = ctx->bvh_valid ? ctx->bvh_prims : nullptr;
    sc.t = terrain_layout(settings.globals.ground_counts[2]);
    const bool profiled = ctx->ring_profile && ctx->ring_profile_body < settings.counts[0];
    sc.ring_profile = profiled ? ctx->ring_profile : nullptr;
    sc.ring_profile_body = profiled ? ctx->ring_profile_body : SG_MISS;
    sc.density_profile = ctx->density_on ? ctx->density_profile : nullptr;
    return sc;
}

// History key of a frame (src/pathtracer.rs:172-181): everything whose change
// restarts the accumulation.
std::vector<uint8_t> history_key(const sg_ctx *ctx, const sg_frame *frame) {
    std::vector<uint8_t> key;
    const uint8_t *g = reinterpret_cast<const uint8_t *>(frame->globals);
    key.insert(key.end(), g, g + sizeof(sg_globals));
    const uint8_t *b = reinterpret_cast<const uint8_t *>(frame->bodies);
    if (frame->body_count) key.insert(key.end(), b, b + sizeof(sg_sphere) * frame->body_count);
    const uint8_t *p = reinterpret_cast<const uint8_t *>(frame->ground);
    if (frame->ground_count) key.insert(key.end(), p, p + sizeof(sg_primitive) * frame->ground_count);
    const uint8_t *t = reinterpret_cast<const uint8_t *>(&frame->scene_time);
    key.insert(key.end(), t, t + 8);
    const uint8_t *mb = reinterpret_cast<const uint8_t *>(&ctx->options.max_bounces);
    key.insert(key.end(), mb, mb + 4);
    const uint8_t *mw = reinterpret_cast<const uint8_t *>(&ctx->options.milky_way);
    key.insert(key.end(), mw, mw + 4);
    return key;
}

// History::update (src/pathtracer.rs:172-188).
void history_update(sg_ctx *ctx, const sg_frame *frame) {
    std::vector<uint8_t> key = history_key(ctx, frame);
    if (key != ctx->key) {
        ctx->key.swap(key);
        ctx->samples = 0;
        ctx->frame_on_device = false;
    }
}

int validate_frame(sg_ctx *ctx, const sg_frame *frame) {
    if (!frame || !frame->globals) return fail(ctx, SG_E_INVALID, "frame and frame->globals are required");
    if (frame->body_count > SG_BODY_CAPACITY) return fail(ctx, SG_E_CAPACITY, "too many scene bodies");
    if (frame->body_count && !frame->bodies) return fail(ctx, SG_E_INVALID, "bodies pointer is null");
    if (frame->ground_count && !frame->ground) return fail(ctx, SG_E_INVALID, "ground pointer is null");
    if (frame->globals->ground_counts[0] != frame->ground_count)
        return fail(ctx, SG_E_INVALID, "globals.ground_counts[0] must equal the primitive count");
    for (uint32_t i = 0; i < frame->body_count; ++i) {
        const float style = frame->bodies[i].axis_x[3];
        if (frame->bodies[i].emissive < 0.5f && (uint32_t)(style + 0.5f) > 1u)
            return fail(ctx, SG_E_INVALID, "surface styles 2-6 (cratered, rocky, icy, volcanic, terran) are not built");
    }
    const sg_globals *g = frame->globals;
    if (g->ground_north[3] != 0.0f) {
        if ((g->ground_counts[2] & 255u) != 0u)
            return fail(ctx, SG_E_INVALID, "only ground theme 0 (forest) is built");
        uint32_t host = (uint32_t)g->ground_north[3] - 1u;
        if (host >= frame->body_count) return fail(ctx, SG_E_INVALID, "ground host index out of range");
        if (g->ground_counts[1] != 0u && !ctx->terrain)
            return fail(ctx, SG_E_STATE, "terrain enabled but no terrain uploaded (sg_upload_terrain)");
    }
    if (g->atmo_params[3] >= 0.5f && (uint32_t)g->atmo_params[3] - 1u >= frame->body_count)
        return fail(ctx, SG_E_INVALID, "atmosphere host index out of range");
    return SG_OK;
}

int ensure_objects(sg_ctx *ctx, uint32_t count) {
    if (count <= ctx->object_capacity && ctx->objects) return SG_OK;
    uint32_t cap = std::max<uint32_t>(SG_GROUND_CAPACITY, count);
    if (ctx->objects) SG_HIP(ctx, hipFree(ctx->objects));
    ctx->objects = nullptr;
    SG_HIP(ctx, hipMalloc(&ctx->objects, (size_t)cap * sizeof(sg_primitive)));
    ctx->object_capacity = cap;
    return SG_OK;
}

int ensure_wave(sg_ctx *ctx, uint32_t paths) {
    if (paths <= ctx->wave_capacity && ctx->wave_block) return SG_OK;
    if (ctx->wave_block) SG_HIP(ctx, hipFree(ctx->wave_block));
    if (ctx->queues[0]) SG_HIP(ctx, hipFree(ctx->queues[0]));
    ctx->wave_block = nullptr;
    ctx->queues[0] = nullptr;
    // ray_o, prev_o: 32 B; ray_d, thr, rad, rbase, h0, h1, h2, value, backdrop: 16 B; rng: 4 B.
    const size_t n = paths;
    const size_t bytes = n * (32 * 2 + 16 * 9 + 4) + 256 * 12;
    SG_HIP(ctx, hipMalloc(&ctx->wave_block, bytes));
    char *p = static_cast<char *>(ctx->wave_block);
    auto take = [&](size_t each) {
        char *q = p;
        p += (n * each + 255) & ~size_t(255);
        return q;
    };
    WaveArgs &w = ctx->wave;
    w.ray_o = reinterpret_cast<double4 *>(take(32));
    w.prev_o = reinterpret_cast<double4 *>(take(32));
    w.ray_d = reinterpret_cast<float4 *>(take(16));
    w.thr = reinterpret_cast<float4 *>(take(16));
    w.rad = reinterpret_cast<float4 *>(take(16));
    w.rbase = reinterpret_cast<float4 *>(take(16));
    w.h0 = reinterpret_cast<float4 *>(take(16));
    w.h1 = reinterpret_cast<float4 *>(take(16));
    w.h2 = reinterpret_cast<float4 *>(take(16));
    w.value = reinterpret_cast<float4 *>(take(16));
    w.backdrop = reinterpret_cast<float4 *>(take(16));
    w.rng = reinterpret_cast<uint32_t *>(take(4));
    SG_HIP(ctx, hipMalloc(&ctx->queues[0], (2 * n + 64) * sizeof(uint32_t)));
    ctx->queues[1] = ctx->queues[0] + n;
    ctx->counter = ctx->queues[0] + 2 * n;
    if (!ctx->counter_host) SG_HIP(ctx, hipHostMalloc(&ctx->counter_host, 64, hipHostMallocDefault));
    ctx->wave_capacity = paths;
    return SG_OK;
}

// One progressive pass through the wavefront kernels, chunked over pixels so
// the path state stays within wave_capacity.
int encode_wavefront(sg_ctx *ctx, const Scene &sc, const TraceArgs &ta) {
    const uint32_t count = ta.count;
    const uint64_t pixels = (uint64_t)ta.width * ta.height;
    const uint32_t chunk_pixels = std::max<uint32_t>(1u, ctx->chunk_paths / count);
    const uint32_t capacity = (uint32_t)std::min<uint64_t>(pixels, chunk_pixels) * count;
    int rc = ensure_wave(ctx, capacity);
    if (rc != SG_OK) return rc;
    WaveArgs wa = ctx->wave;
    wa.width = ta.width;
    wa.height = ta.height;
    wa.count = count;
    wa.accumulated = ta.accumulated;
    wa.sequence = ta.sequence;
    wa.has_air = ta.has_air;
    wa.records = ta.records;
    const uint32_t block = 256;
    for (uint64_t p0 = 0; p0 < pixels; p0 += chunk_pixels) {
        const uint32_t chunk = (uint32_t)std::min<uint64_t>(chunk_pixels, pixels - p0);
        wa.pixel_begin = (uint32_t)p0;
        wa.paths = chunk * count;
        uint32_t n = wa.paths;
        int q = 0;
        hipLaunchKernelGGL(k_generate, dim3((n + block - 1) / block), dim3(block), 0, ctx->stream, sc, wa,
                           ctx->queues[q]);
        SG_HIP(ctx, hipGetLastError());
        for (uint32_t iteration = 0; n > 0; ++iteration) {
            if (n < ctx->tail_paths) {
                hipLaunchKernelGGL(k_tail, dim3((n + block - 1) / block), dim3(block), 0, ctx->stream, sc, wa,
                                   ctx->queues[q], n);
                SG_HIP(ctx, hipGetLastError());
                break;
            }
            const dim3 grid((n + block - 1) / block);
            hipLaunchKernelGGL(k_extend, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n);
            SG_HIP(ctx, hipGetLastError());
            if (iteration == 0 && ta.has_air) {
                hipLaunchKernelGGL(k_camera_air, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n);
                SG_HIP(ctx, hipGetLastError());
            }
            SG_HIP(ctx, hipMemsetAsync(ctx->counter, 0, sizeof(uint32_t), ctx->stream));
            hipLaunchKernelGGL(k_shade, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n,
                               ctx->queues[q ^ 1], ctx->counter);
            SG_HIP(ctx, hipGetLastError());
            SG_HIP(ctx, hipMemcpyAsync(ctx->counter_host, ctx->counter, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                       ctx->stream));
            SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            n = *ctx->counter_host;
            q ^= 1;
        }
        hipLaunchKernelGGL(k_accumulate, dim3((chunk + block - 1) / block), dim3(block), 0, ctx->stream, wa,
                           ta.accumulation, ta.backdrop, ta.moments, chunk);
        SG_HIP(ctx, hipGetLastError());
    }
    return SG_OK;
}

// The host half of PathTracer::encode (src/pathtracer.rs:789-960): history,
// sample count, the Settings uniform, uploads and the cache kernels, queued
// on the context stream. Leaves the integrator launch to the caller.
struct PassPlan {
    Scene sc;
    TraceArgs ta;
    uint32_t count = 0;
    bool queued = false;  // work was queued on the context stream
};

// The Settings uniform of a pass of `count` samples over `samples` already
// accumulated, at noise sequence `sequence` (src/pathtracer.rs:829-905).
sg_settings pass_settings(const sg_ctx *ctx, const sg_frame *frame, uint32_t count, uint32_t samples,
                          uint32_t sequence) {
    const sg_options &o = ctx->options;
    sg_settings st;
    std::memset(&st, 0, sizeof(st));
    st.globals = *frame->globals;
    sg_globals &g = st.globals;
    g.viewport[3] = g.viewport[0] / std::max(g.viewport[1], 1.0f);
    g.viewport[0] = (float)ctx->width;
    g.viewport[1] = (float)ctx->height;
    g.viewport[2] = o.exposure;
    float prop_top = 0.0f;
    for (uint32_t i = 0; i < frame->ground_count; ++i) {
        const sg_primitive &p = frame->ground[i];
        float e = std::sqrt(p.extent[0] * p.extent[0] + p.extent[1] * p.extent[1] + p.extent[2] * p.extent[2]);
        prop_top = std::max(prop_top, p.center[1] + e);
    }
    g.ground_counts[3] = float_bits(prop_top);
    uint32_t cells = ctx->terrain_cells ? ctx->terrain_cells : 512u;
    uint32_t log2 = 0;
    while ((1u << log2) < cells) ++log2;
    g.ground_counts[2] |= log2 << 8;
    uint32_t lights[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < frame->body_count; ++i) {
        if (frame->bodies[i].emissive >= 0.5f) {
            if (lights[3] < 3u) lights[lights[3]] = i;
            lights[3] += 1u;
        }
    }
    std::memcpy(st.lights, lights, sizeof(lights));
    st.sampling[0] = sequence;
    st.counts[0] = frame->body_count;
    st.counts[1] = samples;
    st.counts[2] = count;
    st.counts[3] = o.max_bounces;
    st.background[0] = o.milky_way;
    st.background[1] = ctx->milky_way ? (float)ctx->milky_w : 1024.0f;
    st.background[2] = ctx->milky_way ? (float)ctx->milky_h : 512.0f;
    st.background[3] = 100.0f;  // stars::POINT_SOURCE_GAIN (src/stars.rs:130)
    return st;
}

int prepare_pass(sg_ctx *ctx, const sg_frame *frame, PassPlan &plan) {
    int rc = validate_frame(ctx, frame);
    if (rc != SG_OK) return rc;
    if (frame->ground_count > 0) {
        rc = ensure_objects(ctx, frame->ground_count);
        if (rc != SG_OK) return rc;
    }
    history_update(ctx, frame);
    const sg_options &o = ctx->options;
    uint32_t remaining = o.sample_limit > ctx->samples ? o.sample_limit - ctx->samples : 0u;
    uint32_t count = std::min(std::min(o.samples_per_frame, SG_MAX_BATCH_SAMPLES), remaining);
    sg_settings st = pass_settings(ctx, frame, count, ctx->samples, o.noise_seed + ctx->sequence);
    const sg_globals &g = st.globals;
    const bool has_air = frame->globals->atmo_params[3] >= 0.5f;
    plan.count = count;
    plan.queued = false;
    if (count == 0) return SG_OK;
    // The device copies of bodies and primitives stay valid while the
    // history key (which covers them) is unchanged.
    if (!ctx->frame_on_device) {
        if (frame->body_count)
            SG_HIP(ctx, hipMemcpyAsync(ctx->bodies, frame->bodies, frame->body_count * sizeof(Body),
                                       hipMemcpyHostToDevice, ctx->stream));
        if (frame->ground_count)
            SG_HIP(ctx, hipMemcpyAsync(ctx->objects, frame->ground, frame->ground_count * sizeof(sg_primitive),
                                       hipMemcpyHostToDevice, ctx->stream));
        ctx->frame_on_device = true;
        plan.queued = true;
    }
    int brc = ensure_bvh(ctx, frame);
    if (brc != SG_OK) return brc;
    plan.sc = make_scene(ctx, st);
    const Scene &sc = plan.sc;
    const bool terrain_on = g.ground_counts[1] != 0u && g.ground_north[3] != 0.0f;
    const uint32_t radius_bits = float_bits(g.ground_up[3]);
    if (terrain_on && (!ctx->terrain_cached || ctx->terrain_radius_bits != radius_bits)) {
        const uint64_t n = sc.t.vertices;
        hipLaunchKernelGGL(k_terrain_cache, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, sc,
                           ctx->terrain);
        SG_HIP(ctx, hipGetLastError());
        ctx->terrain_cached = true;
        ctx->terrain_radius_bits = radius_bits;
        plan.queued = true;
    }
    if (ctx->samples == 0 && has_air && terrain_on) {
        hipLaunchKernelGGL(k_sky_cache, dim3(1), dim3(128), 0, ctx->stream, sc, ctx->sky_light);
        SG_HIP(ctx, hipGetLastError());
        plan.queued = true;
    }
    TraceArgs &ta = plan.ta;
    ta.accumulation = ctx->accumulation;
    ta.backdrop = ctx->backdrop;
    ta.moments = ctx->moments_on ? ctx->moments : nullptr;
    ta.records = nullptr;
    if (ctx->record_on) {
        const uint64_t need = (uint64_t)ctx->width * ctx->height * count;
        if (need > ctx->record_capacity) {
            SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (ctx->records) SG_HIP(ctx, hipFree(ctx->records));
            ctx->records = nullptr;
            ctx->record_capacity = 0;
            SG_HIP(ctx, hipMalloc(&ctx->records, need * sizeof(float4)));
            ctx->record_capacity = need;
        }
        ta.records = ctx->records;
    }
    ta.width = ctx->width;
    ta.height = ctx->height;
    ta.count = count;
    ta.accumulated = ctx->samples;
    ta.sequence = st.sampling[0];
    ta.has_air = has_air ? 1u : 0u;
    ta.pixels_per_block = std::max<uint32_t>(1u, kTraceBlock / count);
    ctx->last.count = count;
    ctx->last.samples = ctx->samples;
    ctx->last.sequence = st.sampling[0];
    ctx->last.recorded = ta.records != nullptr;
    return SG_OK;
}

// The history bookkeeping after a pass (src/pathtracer.rs:1002-1008).
void finish_pass(sg_ctx *ctx, uint32_t count) {
    const sg_options &o = ctx->options;
    ctx->samples += count;
    ctx->sequence += count;
    ctx->exposure_idle_frames = count > 0 ? 0u : std::min<uint32_t>(ctx->exposure_idle_frames + 1u, 0xffffffffu);
    ctx->exposure_was_enabled = o.auto_exposure != 0;
}

}  // namespace

extern "C" {

int sg_create(sg_ctx **out, uint32_t width, uint32_t height, const sg_ray_star *catalogue, uint32_t catalogue_count,
              const uint32_t *sky_cells, uint32_t sky_cell_count, const sg_options *options) {
    if (!out) return SG_E_INVALID;
    *out = nullptr;
    if (width == 0 || height == 0) return SG_E_INVALID;
    sg_ctx *ctx = new sg_ctx();
    ctx->width = width;
    ctx->height = height;
    if (options) {
        ctx->options = *options;
    } else {
        // Options::default (src/pathtracer.rs:97-112)
        ctx->options = sg_options{0u, 1u, 8u, SG_MAX_SAMPLES, 1.0f, 0, 0.0f, 0.25f, 0.9f, 4.0e-4f};
    }
    auto bail = [&](int code) {
        sg_destroy(ctx);
        return code;
    };
    if (hipGetDevice(&ctx->device) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_ready, hipEventDisableTiming) != hipSuccess)
        return bail(SG_E_DEVICE);
    if (alloc_image(ctx) != SG_OK) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->bodies, SG_BODY_CAPACITY * sizeof(Body)) != hipSuccess) return bail(SG_E_DEVICE);
    if (ensure_objects(ctx, SG_GROUND_CAPACITY) != SG_OK) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->sky_light, SG_SKY_TEXELS * sizeof(float4)) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipMemsetAsync(ctx->sky_light, 0, SG_SKY_TEXELS * sizeof(float4), ctx->stream) != hipSuccess)
        return bail(SG_E_DEVICE);
    if (catalogue_count > 0) {
        const uint32_t cells = 256u * 128u + 1u;
        if (!catalogue || !sky_cells || sky_cell_count < cells) return bail(SG_E_INVALID);
        // Offsets must be monotone, end at the index-list length, and every
        // index must name a catalogue star, so the miss shader stays in bounds.
        for (uint32_t i = 0; i + 1 < cells; ++i)
            if (sky_cells[i] > sky_cells[i + 1]) return bail(SG_E_INVALID);
        if ((uint64_t)sky_cells[cells - 1] + cells != sky_cell_count) return bail(SG_E_INVALID);
        for (uint32_t i = cells; i < sky_cell_count; ++i)
            if (sky_cells[i] >= catalogue_count) return bail(SG_E_INVALID);
        if (hipMalloc(&ctx->catalogue, catalogue_count * sizeof(sg_ray_star)) != hipSuccess ||
            hipMalloc(&ctx->sky_cells, sky_cell_count * sizeof(uint32_t)) != hipSuccess)
            return bail(SG_E_DEVICE);
        if (hipMemcpy(ctx->catalogue, catalogue, catalogue_count * sizeof(sg_ray_star), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(ctx->sky_cells, sky_cells, sky_cell_count * sizeof(uint32_t), hipMemcpyHostToDevice) !=
                hipSuccess)
            return bail(SG_E_DEVICE);
        ctx->catalogue_count = catalogue_count;
    }
    *out = ctx;
    return SG_OK;
}

int sg_destroy(sg_ctx *ctx) {
    if (!ctx) return SG_OK;
    if (ctx->batch) return fail(ctx, SG_E_STATE, "context is a view of a batch; destroy the batch first");
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    void *buffers[] = {ctx->accumulation, ctx->backdrop, ctx->moments, ctx->bodies, ctx->objects, ctx->sky_light,
                       ctx->terrain, ctx->catalogue, ctx->sky_cells, ctx->milky_way, ctx->probe_rays,
                       ctx->probe_results, ctx->wave_block, ctx->queues[0], ctx->bvh_nodes,
                       ctx->bvh_prims, ctx->ring_profile, ctx->density_profile, ctx->adj_image,
                       ctx->adj_gradient, ctx->adj_primal, ctx->records};
    if (ctx->counter_host) (void)hipHostFree(ctx->counter_host);
    for (void *b : buffers)
        if (b) (void)hipFree(b);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_ready) (void)hipEventDestroy(ctx->ev_ready);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SG_OK;
}

int sg_set_milky_way(sg_ctx *ctx, const uint32_t *texels, uint32_t width, uint32_t height) {
    if (!ctx || !texels || width == 0 || height == 0) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->milky_way) SG_HIP(ctx, hipFree(ctx->milky_way));
    ctx->milky_way = nullptr;
    SG_HIP(ctx, hipMalloc(&ctx->milky_way, (size_t)width * height * 4u));
    SG_HIP(ctx, hipMemcpy(ctx->milky_way, texels, (size_t)width * height * 4u, hipMemcpyHostToDevice));
    ctx->milky_w = width;
    ctx->milky_h = height;
    return SG_OK;
}

int sg_resize(sg_ctx *ctx, uint32_t width, uint32_t height) {
    if (!ctx || width == 0 || height == 0) return SG_E_INVALID;
    if (width == ctx->width && height == ctx->height) return SG_OK;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->width = width;
    ctx->height = height;
    int rc = alloc_image(ctx);
    if (rc != SG_OK) return rc;
    ctx->key.clear();
    ctx->samples = 0;
    ctx->exposure_idle_frames = 0;
    return SG_OK;
}

uint32_t sg_samples(const sg_ctx *ctx) { return ctx ? ctx->samples : 0u; }

int sg_size(const sg_ctx *ctx, uint32_t *width, uint32_t *height) {
    if (!ctx || !width || !height) return SG_E_INVALID;
    *width = ctx->width;
    *height = ctx->height;
    return SG_OK;
}

int sg_needs_redraw(const sg_ctx *ctx) {
    if (!ctx) return SG_E_INVALID;
    return (ctx->samples < ctx->options.sample_limit || (ctx->options.auto_exposure && ctx->exposure_idle_frames < 60))
               ? 1
               : 0;
}

int sg_prepare_frame(sg_ctx *ctx, const sg_frame *frame) {
    if (!ctx || !frame || !frame->globals) return SG_E_INVALID;
    history_update(ctx, frame);
    return ctx->samples == 0 ? 1 : 0;
}

int sg_set_quality(sg_ctx *ctx, uint32_t samples, uint32_t bounces, uint32_t limit) {
    if (!ctx) return SG_E_INVALID;
    bounces = std::min<uint32_t>(std::max<uint32_t>(bounces, 1u), 64u);
    if (ctx->options.max_bounces != bounces) {
        ctx->key.clear();
        ctx->samples = 0;
        ctx->exposure_idle_frames = 0;
    }
    ctx->options.samples_per_frame = std::min<uint32_t>(std::max<uint32_t>(samples, 1u), 64u);
    ctx->options.max_bounces = bounces;
    ctx->options.sample_limit = std::min<uint32_t>(std::max<uint32_t>(limit, 1u), SG_MAX_SAMPLES);
    return SG_OK;
}

int sg_set_exposure_controls(sg_ctx *ctx, int32_t auto_exposure, float bias) {
    if (!ctx) return SG_E_INVALID;
    ctx->options.auto_exposure = auto_exposure ? 1 : 0;
    ctx->options.ev_bias = bias;
    ctx->exposure_idle_frames = 0;
    return SG_OK;
}

int sg_upload_terrain(sg_ctx *ctx, const float *data, uint64_t float_count, uint32_t cells) {
    if (!ctx || !data) return SG_E_INVALID;
    if (cells < 4u || cells > SG_MAX_TERRAIN_CELLS || (cells & (cells - 1u)) != 0u)
        return fail(ctx, SG_E_INVALID, "terrain cells must be a power of two in [4, 8192]");
    uint32_t log2 = 0;
    while ((1u << log2) < cells) ++log2;
    TerrainLayout t = terrain_layout(log2 << 8);
    if (float_count != t.data) return fail(ctx, SG_E_INVALID, "terrain data size does not match its cell count");
    const uint64_t total = t.total;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->terrain && ctx->terrain_floats != total) {
        SG_HIP(ctx, hipFree(ctx->terrain));
        ctx->terrain = nullptr;
    }
    if (!ctx->terrain) SG_HIP(ctx, hipMalloc(&ctx->terrain, total * sizeof(float)));
    SG_HIP(ctx, hipMemcpy(ctx->terrain, data, float_count * sizeof(float), hipMemcpyHostToDevice));
    ctx->terrain_floats = total;
    ctx->terrain_cells = cells;
    ctx->terrain_cached = false;
    return SG_OK;
}

int sg_enable_moments(sg_ctx *ctx, int32_t enable) {
    if (!ctx) return SG_E_INVALID;
    bool on = enable != 0;
    if (on == ctx->moments_on) return SG_OK;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->moments_on = on;
    ctx->key.clear();
    ctx->samples = 0;
    return alloc_image(ctx);
}

// PathTracer::encode (src/pathtracer.rs:789-1009).
int sg_encode(sg_ctx *ctx, const sg_frame *frame) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    PassPlan plan;
    int rc = prepare_pass(ctx, frame, plan);
    if (rc != SG_OK) return rc;
    if (plan.count > 0) {
        const TraceArgs &ta = plan.ta;
        const uint32_t block = ta.pixels_per_block * plan.count;
        const uint64_t pixels = (uint64_t)ctx->width * ctx->height;
        const uint32_t grid = (uint32_t)((pixels + ta.pixels_per_block - 1) / ta.pixels_per_block);
        if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
        if (ctx->integrator == SG_INTEGRATOR_FUSED) {
            hipLaunchKernelGGL(k_trace, dim3(grid), dim3(block), 0, ctx->stream, plan.sc, ta);
            SG_HIP(ctx, hipGetLastError());
        } else {
            int wrc = encode_wavefront(ctx, plan.sc, ta);
            if (wrc != SG_OK) return wrc;
        }
        if (ctx->profiling) {
            SG_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
            SG_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
            float ms = 0.f;
            SG_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
            ctx->kernel_ms += ms;
            ctx->launches += 1;
        }
    }
    finish_pass(ctx, plan.count);
    return SG_OK;
}

static int read_image(sg_ctx *ctx, const float4 *src, float *dst, uint64_t pixel_count) {
    if (!ctx || !dst) return SG_E_INVALID;
    if (!src) return fail(ctx, SG_E_STATE, "buffer not allocated");
    if (pixel_count != (uint64_t)ctx->width * ctx->height)
        return fail(ctx, SG_E_INVALID, "pixel count does not match the image size");
    join_batch(ctx);
    SG_HIP(ctx, hipMemcpyAsync(dst, src, pixel_count * 16u, hipMemcpyDeviceToHost, ctx->stream));
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}
int sg_read_accumulation(sg_ctx *ctx, float *rgba, uint64_t n) { return read_image(ctx, ctx ? ctx->accumulation : nullptr, rgba, n); }
int sg_read_backdrop(sg_ctx *ctx, float *rgba, uint64_t n) { return read_image(ctx, ctx ? ctx->backdrop : nullptr, rgba, n); }
int sg_read_moments(sg_ctx *ctx, float *rgba, uint64_t n) { return read_image(ctx, ctx ? ctx->moments : nullptr, rgba, n); }

int sg_synchronize(sg_ctx *ctx) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

int sg_set_profiling(sg_ctx *ctx, int32_t enable) {
    if (!ctx) return SG_E_INVALID;
    ctx->profiling = enable != 0;
    ctx->kernel_ms = 0.0;
    ctx->launches = 0;
    ctx->adjoint_ms = 0.0;
    ctx->adjoint_launches = 0;
    return SG_OK;
}

int sg_kernel_time(sg_ctx *ctx, double *milliseconds, uint64_t *launches) {
    if (!ctx || !milliseconds || !launches) return SG_E_INVALID;
    *milliseconds = ctx->kernel_ms;
    *launches = ctx->launches;
    return SG_OK;
}

int sg_set_integrator(sg_ctx *ctx, uint32_t integrator, uint32_t chunk_paths) {
    if (!ctx) return SG_E_INVALID;
    if (integrator != SG_INTEGRATOR_FUSED && integrator != SG_INTEGRATOR_WAVEFRONT)
        return fail(ctx, SG_E_INVALID, "unknown integrator");
    if (chunk_paths != 0 && (chunk_paths < 64u || chunk_paths > (1u << 28)))
        return fail(ctx, SG_E_INVALID, "chunk_paths must be 0 (keep) or in [64, 2^28]");
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->integrator = integrator;
    if (chunk_paths) ctx->chunk_paths = chunk_paths;
    return SG_OK;
}

int sg_set_tail(sg_ctx *ctx, uint32_t tail_paths) {
    if (!ctx) return SG_E_INVALID;
    ctx->tail_paths = tail_paths;
    return SG_OK;
}

int sg_get_launch(const sg_ctx *ctx, sg_launch *launch) {
    if (!ctx || !launch) return SG_E_INVALID;
    *launch = sg_launch{ctx->integrator, ctx->chunk_paths, ctx->tail_paths};
    return SG_OK;
}

int sg_autotune(sg_ctx *ctx, const sg_frame *frame, uint32_t repeats, sg_launch *best, double *best_ms) {
    if (!ctx || !frame) return SG_E_INVALID;
    repeats = std::max<uint32_t>(1u, std::min<uint32_t>(repeats, 16u));
    const sg_launch keep{ctx->integrator, ctx->chunk_paths, ctx->tail_paths};
    const uint32_t keep_sequence = ctx->sequence;
    const bool keep_profiling = ctx->profiling;
    const double keep_ms = ctx->kernel_ms;
    const uint64_t keep_launches = ctx->launches;
    sg_launch winner = keep;
    double winner_ms = 1e300;
    int rc = SG_OK;
    auto trial = [&](const sg_launch &l) -> int {
        ctx->integrator = l.integrator;
        ctx->chunk_paths = l.chunk_paths;
        ctx->tail_paths = l.tail_paths;
        int r = sg_encode(ctx, frame);  // warm: buffers sized, caches filled
        if (r != SG_OK) return r;
        ctx->profiling = true;
        ctx->kernel_ms = 0.0;
        ctx->launches = 0;
        for (uint32_t i = 0; i < repeats && r == SG_OK; ++i) r = sg_encode(ctx, frame);
        ctx->profiling = false;
        if (r != SG_OK) return r;
        if (ctx->launches == 0) return SG_OK;  // sample limit reached: nothing to time
        const double ms = ctx->kernel_ms / (double)ctx->launches;
        if (ms < winner_ms) {
            winner_ms = ms;
            winner = l;
        }
        return SG_OK;
    };
    // Queue capacity first (no tail), then the compaction threshold at the
    // best capacity, then the fused integrator.
    const uint32_t chunks[] = {1u << 21, 1u << 22, 1u << 23, 1u << 24};
    for (uint32_t c : chunks) {
        rc = trial(sg_launch{SG_INTEGRATOR_WAVEFRONT, c, 0u});
        if (rc != SG_OK) break;
    }
    if (rc == SG_OK) {
        const uint32_t capacity = winner.chunk_paths;
        const uint32_t tails[] = {capacity / 64u, capacity / 16u, capacity / 4u};
        for (uint32_t t : tails) {
            rc = trial(sg_launch{SG_INTEGRATOR_WAVEFRONT, capacity, t});
            if (rc != SG_OK) break;
        }
    }
    if (rc == SG_OK) rc = trial(sg_launch{SG_INTEGRATOR_FUSED, keep.chunk_paths, keep.tail_paths});
    if (rc == SG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, SG_E_DEVICE, "autotune sync");
    const sg_launch chosen = (rc == SG_OK && winner_ms < 1e300) ? winner : keep;
    ctx->integrator = chosen.integrator;
    ctx->chunk_paths = chosen.chunk_paths;
    ctx->tail_paths = chosen.tail_paths;
    ctx->profiling = keep_profiling;
    ctx->kernel_ms = keep_ms;
    ctx->launches = keep_launches;
    // Trial passes rendered into the image: start the next pass afresh, with
    // the draws it would have had without them.
    ctx->key.clear();
    ctx->samples = 0;
    ctx->sequence = keep_sequence;
    ctx->last.count = 0;
    if (rc != SG_OK) return rc;
    if (best) *best = chosen;
    if (best_ms) *best_ms = winner_ms;
    return SG_OK;
}

const char *sg_last_error(const sg_ctx *ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int sg_enable_adjoint(sg_ctx *ctx, int32_t enable) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->record_on = enable != 0;
    ctx->last.recorded = false;
    if (!ctx->record_on && ctx->records) {
        SG_HIP(ctx, hipFree(ctx->records));
        ctx->records = nullptr;
        ctx->record_capacity = 0;
    }
    return SG_OK;
}

int sg_restart(sg_ctx *ctx) {
    if (!ctx) return SG_E_INVALID;
    ctx->key.clear();
    ctx->samples = 0;
    ctx->last.count = 0;
    ctx->exposure_idle_frames = 0;
    return SG_OK;
}

int sg_set_ring_profile(sg_ctx *ctx, uint32_t body, const float *tau, uint32_t bins) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!tau) {
        ctx->ring_profile_body = SG_MISS;
    } else {
        if (bins != SG_RING_PROFILE_BINS) return fail(ctx, SG_E_INVALID, "the ring profile has SG_RING_PROFILE_BINS bins");
        if (body >= SG_BODY_CAPACITY) return fail(ctx, SG_E_INVALID, "ring profile body index out of range");
        for (uint32_t i = 0; i < bins; ++i)
            if (!(tau[i] >= 0.0f) || !std::isfinite(tau[i]))
                return fail(ctx, SG_E_INVALID, "ring optical depths must be finite and non-negative");
        if (!ctx->ring_profile) SG_HIP(ctx, hipMalloc(&ctx->ring_profile, SG_RING_PROFILE_BINS * sizeof(float)));
        SG_HIP(ctx, hipMemcpy(ctx->ring_profile, tau, bins * sizeof(float), hipMemcpyHostToDevice));
        ctx->ring_profile_body = body;
    }
    // A new profile is a new scene: restart the accumulation.
    ctx->key.clear();
    ctx->samples = 0;
    ctx->last.count = 0;
    return SG_OK;
}

int sg_set_density_profile(sg_ctx *ctx, const float *density, uint32_t bins) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!density) {
        ctx->density_on = false;
    } else {
        if (bins != SG_DENSITY_BINS) return fail(ctx, SG_E_INVALID, "the density profile has SG_DENSITY_BINS bins");
        for (uint32_t i = 0; i < bins; ++i)
            if (!(density[i] >= 0.0f) || !std::isfinite(density[i]))
                return fail(ctx, SG_E_INVALID, "air densities must be finite and non-negative");
        if (!ctx->density_profile) SG_HIP(ctx, hipMalloc(&ctx->density_profile, SG_DENSITY_BINS * sizeof(float)));
        SG_HIP(ctx, hipMemcpy(ctx->density_profile, density, bins * sizeof(float), hipMemcpyHostToDevice));
        ctx->density_on = true;
    }
    ctx->key.clear();
    ctx->samples = 0;
    ctx->last.count = 0;
    return SG_OK;
}

int sg_adjoint(sg_ctx *ctx, const sg_frame *frame, const float *d_image, uint64_t pixel_count, float *gradient,
               uint32_t gradient_count, float *pass_mean) {
    if (!ctx || !frame || !frame->globals || !d_image || !gradient) return SG_E_INVALID;
    if (gradient_count != SG_GRAD_COUNT) return fail(ctx, SG_E_INVALID, "gradient_count must be SG_GRAD_COUNT");
    if (pixel_count != (uint64_t)ctx->width * ctx->height)
        return fail(ctx, SG_E_INVALID, "pixel count does not match the image size");
    int rc = validate_frame(ctx, frame);
    if (rc != SG_OK) return rc;
    if (frame->globals->ground_north[3] != 0.0f)
        return fail(ctx, SG_E_INVALID, "the adjoint does not differentiate scenes with a landscape");
    if (ctx->last.count == 0 || ctx->last.samples + ctx->last.count != ctx->samples || history_key(ctx, frame) != ctx->key)
        return fail(ctx, SG_E_STATE, "sg_adjoint needs the frame of the last pass sg_encode rendered");
    join_batch(ctx);
    const uint64_t pixels = pixel_count;
    if (pixels > ctx->adj_pixels) {
        SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->adj_image) SG_HIP(ctx, hipFree(ctx->adj_image));
        if (ctx->adj_primal) SG_HIP(ctx, hipFree(ctx->adj_primal));
        ctx->adj_image = nullptr;
        ctx->adj_primal = nullptr;
        SG_HIP(ctx, hipMalloc(&ctx->adj_image, pixels * 3u * sizeof(float)));
        SG_HIP(ctx, hipMalloc(&ctx->adj_primal, pixels * 2u * sizeof(float4)));
        ctx->adj_pixels = pixels;
    }
    if (!ctx->adj_gradient) SG_HIP(ctx, hipMalloc(&ctx->adj_gradient, SG_GRAD_COUNT * sizeof(float)));
    if (!ctx->frame_on_device) {
        if (frame->body_count)
            SG_HIP(ctx, hipMemcpyAsync(ctx->bodies, frame->bodies, frame->body_count * sizeof(Body),
                                       hipMemcpyHostToDevice, ctx->stream));
        if (frame->ground_count)
            SG_HIP(ctx, hipMemcpyAsync(ctx->objects, frame->ground, frame->ground_count * sizeof(sg_primitive),
                                       hipMemcpyHostToDevice, ctx->stream));
        ctx->frame_on_device = true;
    }
    rc = ensure_bvh(ctx, frame);
    if (rc != SG_OK) return rc;
    const uint32_t count = ctx->last.count;
    Scene sc = make_scene(ctx, pass_settings(ctx, frame, count, ctx->last.samples, ctx->last.sequence));
    AdjointArgs aa;
    TraceArgs &ta = aa.ta;
    ta.accumulation = pass_mean ? ctx->adj_primal : nullptr;
    ta.backdrop = pass_mean ? ctx->adj_primal + pixels : nullptr;
    ta.moments = nullptr;
    ta.records = nullptr;
    ta.width = ctx->width;
    ta.height = ctx->height;
    ta.count = count;
    ta.accumulated = ctx->last.samples;
    ta.sequence = ctx->last.sequence;
    ta.has_air = frame->globals->atmo_params[3] >= 0.5f ? 1u : 0u;
    ta.pixels_per_block = std::max<uint32_t>(1u, kTraceBlock / count);
    aa.recorded = (ctx->last.recorded && !pass_mean) ? ctx->records : nullptr;
    aa.d_image = ctx->adj_image;
    aa.gradient = ctx->adj_gradient;
    aa.blocks = (uint32_t)((pixels + ta.pixels_per_block - 1) / ta.pixels_per_block);
    SG_HIP(ctx, hipMemcpyAsync(ctx->adj_image, d_image, pixels * 3u * sizeof(float), hipMemcpyHostToDevice,
                               ctx->stream));
    hipLaunchKernelGGL(k_fill, dim3((SG_GRAD_COUNT + 255) / 256), dim3(256), 0, ctx->stream, ctx->adj_gradient,
                       (uint32_t)SG_GRAD_COUNT, 0.0f);
    SG_HIP(ctx, hipGetLastError());
    int cus = 0;
    SG_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const uint32_t grid = std::min<uint32_t>(aa.blocks, 2u * (uint32_t)std::max(cus, 1));
    if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    hipLaunchKernelGGL(k_adjoint, dim3(grid), dim3(ta.pixels_per_block * count), 0, ctx->stream, sc, aa);
    SG_HIP(ctx, hipGetLastError());
    if (ctx->profiling) {
        SG_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
        SG_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
        float ms = 0.f;
        SG_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
        ctx->adjoint_ms += ms;
        ctx->adjoint_launches += 1;
    }
    SG_HIP(ctx, hipMemcpyAsync(gradient, ctx->adj_gradient, SG_GRAD_COUNT * sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
    if (pass_mean)
        SG_HIP(ctx, hipMemcpyAsync(pass_mean, ctx->adj_primal, pixels * sizeof(float4), hipMemcpyDeviceToHost,
                                   ctx->stream));
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

int sg_adjoint_time(sg_ctx *ctx, double *milliseconds, uint64_t *launches) {
    if (!ctx || !milliseconds || !launches) return SG_E_INVALID;
    *milliseconds = ctx->adjoint_ms;
    *launches = ctx->adjoint_launches;
    return SG_OK;
}

int sg_probe(sg_ctx *ctx, const sg_frame *frame, uint32_t kind, const sg_probe_ray *rays, uint32_t count,
             sg_probe_result *results) {
    if (!ctx || !frame || !frame->globals || (count && (!rays || !results))) return SG_E_INVALID;
    if (kind > SG_PROBE_CONE) return fail(ctx, SG_E_INVALID, "unknown probe kind");
    if (frame->body_count > SG_BODY_CAPACITY) return fail(ctx, SG_E_CAPACITY, "too many scene bodies");
    if (count == 0) return SG_OK;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t a = rays[i].anchor;
        if (a != SG_MISS && a >= frame->body_count) return fail(ctx, SG_E_INVALID, "probe anchor out of range");
        if ((kind == SG_PROBE_SPHERE || kind == SG_PROBE_RING || kind == SG_PROBE_VISIBILITY) &&
            rays[i].target >= frame->body_count)
            return fail(ctx, SG_E_INVALID, "probe target out of range");
    }
    if ((kind == SG_PROBE_TERRAIN || frame->globals->ground_counts[1] != 0u) && !ctx->terrain)
        return fail(ctx, SG_E_STATE, "terrain probe without uploaded terrain");
    join_batch(ctx);
    if (frame->ground_count) {
        int rc = ensure_objects(ctx, frame->ground_count);
        if (rc != SG_OK) return rc;
    }
    if (count > ctx->probe_capacity) {
        if (ctx->probe_rays) SG_HIP(ctx, hipFree(ctx->probe_rays));
        if (ctx->probe_results) SG_HIP(ctx, hipFree(ctx->probe_results));
        ctx->probe_rays = nullptr;
        ctx->probe_results = nullptr;
        SG_HIP(ctx, hipMalloc(&ctx->probe_rays, (size_t)count * sizeof(sg_probe_ray)));
        SG_HIP(ctx, hipMalloc(&ctx->probe_results, (size_t)count * sizeof(sg_probe_result)));
        ctx->probe_capacity = count;
    }
    sg_settings st;
    std::memset(&st, 0, sizeof(st));
    st.globals = *frame->globals;
    float prop_top = 0.0f;
    for (uint32_t i = 0; i < frame->ground_count; ++i) {
        const sg_primitive &p = frame->ground[i];
        float e = std::sqrt(p.extent[0] * p.extent[0] + p.extent[1] * p.extent[1] + p.extent[2] * p.extent[2]);
        prop_top = std::max(prop_top, p.center[1] + e);
    }
    st.globals.ground_counts[3] = float_bits(prop_top);
    uint32_t cells = ctx->terrain_cells ? ctx->terrain_cells : 512u;
    uint32_t log2 = 0;
    while ((1u << log2) < cells) ++log2;
    st.globals.ground_counts[2] = (st.globals.ground_counts[2] & 255u) | (log2 << 8);
    st.counts[0] = frame->body_count;
    uint32_t lights[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < frame->body_count; ++i)
        if (frame->bodies[i].emissive >= 0.5f) {
            if (lights[3] < 3u) lights[lights[3]] = i;
            lights[3] += 1u;
        }
    std::memcpy(st.lights, lights, sizeof(lights));
    // The probe's frame replaces the encoded frame's device copies.
    ctx->frame_on_device = false;
    if (frame->body_count)
        SG_HIP(ctx, hipMemcpyAsync(ctx->bodies, frame->bodies, frame->body_count * sizeof(Body), hipMemcpyHostToDevice,
                                   ctx->stream));
    if (frame->ground_count)
        SG_HIP(ctx, hipMemcpyAsync(ctx->objects, frame->ground, frame->ground_count * sizeof(sg_primitive),
                                   hipMemcpyHostToDevice, ctx->stream));
    int brc = ensure_bvh(ctx, frame);
    if (brc != SG_OK) return brc;
    Scene sc = make_scene(ctx, st);
    const uint32_t radius_bits = float_bits(st.globals.ground_up[3]);
    if (ctx->terrain && (!ctx->terrain_cached || ctx->terrain_radius_bits != radius_bits)) {
        hipLaunchKernelGGL(k_terrain_cache, dim3((unsigned)((sc.t.vertices + 255) / 256)), dim3(256), 0, ctx->stream,
                           sc, ctx->terrain);
        SG_HIP(ctx, hipGetLastError());
        ctx->terrain_cached = true;
        ctx->terrain_radius_bits = radius_bits;
    }
    SG_HIP(ctx, hipMemcpyAsync(ctx->probe_rays, rays, (size_t)count * sizeof(sg_probe_ray), hipMemcpyHostToDevice,
                               ctx->stream));
    hipLaunchKernelGGL(k_probe, dim3((count + 63) / 64), dim3(64), 0, ctx->stream, sc, kind, ctx->probe_rays, count,
                       ctx->probe_results);
    SG_HIP(ctx, hipGetLastError());
    SG_HIP(ctx, hipMemcpyAsync(results, ctx->probe_results, (size_t)count * sizeof(sg_probe_result),
                               hipMemcpyDeviceToHost, ctx->stream));
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched views (no reference counterpart: the reference renders one view per
// PathTracer). A batch borrows existing contexts as its views; each pass runs
// every view's host preparation, then one persistent k_batch launch, replayed
// from a hipGraph while the view table stays the same.
// ---------------------------------------------------------------------------
struct sg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<sg_ctx *> views;
    sg_batch_launch launch{};
    uint32_t compute_units = 0;
    double clock_khz = 0.0;
    // Device state: table, counters, completion clocks.
    BatchView *table = nullptr;
    uint32_t *counter = nullptr;
    unsigned long long *done = nullptr;  // entries + 2 clock words
    BatchView *staging = nullptr;        // pinned copy of the table
    std::vector<BatchView> expected;     // device table after the last pass
    std::vector<uint32_t> entry_view;
    uint32_t total_blocks = 0;
    bool table_valid = false;
    bool pass_queued = false;
    // Graph of one pass for the current table shape and launch.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    hipEvent_t ev_done = nullptr, ev_start = nullptr, ev_stop = nullptr;
    bool profiling = false;
    double kernel_ms = 0.0;
    uint64_t launches = 0;
    std::vector<hipEvent_t> view_events;  // per-view launch completion
    std::string error;
};

namespace {

int bfail(sg_batch *b, int code, const std::string &message) {
    if (b) b->error = message;
    return code;
}
#define SG_BHIP(b, call)                                                                         \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return bfail((b), SG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

bool valid_launch(const sg_batch_launch &l) {
    return (l.block == 64u || l.block == 128u || l.block == 256u) && l.workgroups >= 1u && l.workgroups <= 65536u &&
           l.blocks_per_pull >= 1u && l.blocks_per_pull <= 64u && l.order <= SG_BATCH_ORDER_LONGEST &&
           l.use_graph <= 1u;
}

void drop_graph(sg_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

BatchArgs batch_args(const sg_batch *b) {
    BatchArgs ba;
    ba.table = b->table;
    ba.entries = (uint32_t)b->entry_view.size();
    ba.total_blocks = b->total_blocks;
    ba.blocks_per_pull = b->launch.blocks_per_pull;
    ba.counter = b->counter;
    ba.done = b->done;
    ba.clock = b->done + kMaxBatchViews;
    return ba;
}

// Queue one pass (reset, persistent trace, advance) on the batch stream.
int queue_pass(sg_batch *b) {
    const BatchArgs ba = batch_args(b);
    const dim3 small((ba.entries + 255) / 256 + 1);
    hipLaunchKernelGGL(k_batch_reset, small, dim3(256), 0, b->stream, ba);
    SG_BHIP(b, hipGetLastError());
    const dim3 grid(b->launch.workgroups);
    switch (b->launch.block) {
    case 64: hipLaunchKernelGGL(k_batch<64>, grid, dim3(64), 0, b->stream, ba); break;
    case 128: hipLaunchKernelGGL(k_batch<128>, grid, dim3(128), 0, b->stream, ba); break;
    default: hipLaunchKernelGGL(k_batch<256>, grid, dim3(256), 0, b->stream, ba); break;
    }
    SG_BHIP(b, hipGetLastError());
    hipLaunchKernelGGL(k_batch_advance, small, dim3(256), 0, b->stream, ba);
    SG_BHIP(b, hipGetLastError());
    return SG_OK;
}

int capture_pass(sg_batch *b) {
    drop_graph(b);
    SG_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = queue_pass(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != SG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, SG_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    SG_BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_valid = true;
    return SG_OK;
}

// The table a pass of `plans` would use under the current launch settings.
void build_table(const sg_batch *b, std::vector<PassPlan> &plans, std::vector<BatchView> &table,
                 std::vector<uint32_t> &entry_view, uint32_t &total_blocks) {
    table.clear();
    entry_view.clear();
    const uint32_t n = (uint32_t)plans.size();
    std::vector<uint32_t> order;
    std::vector<uint64_t> blocks(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (plans[i].count == 0) continue;
        TraceArgs &ta = plans[i].ta;
        ta.pixels_per_block = std::max<uint32_t>(1u, b->launch.block / ta.count);
        const uint64_t pixels = (uint64_t)ta.width * ta.height;
        blocks[i] = (pixels + ta.pixels_per_block - 1) / ta.pixels_per_block;
        order.push_back(i);
    }
    if (b->launch.order == SG_BATCH_ORDER_SHORTEST)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return blocks[x] < blocks[y]; });
    else if (b->launch.order == SG_BATCH_ORDER_LONGEST)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return blocks[x] > blocks[y]; });
    total_blocks = 0;
    for (uint32_t i : order) {
        BatchView v;
        std::memset(&v, 0, sizeof(v));
        v.scene = plans[i].sc;
        v.ta = plans[i].ta;
        v.block_begin = total_blocks;
        v.blocks = (uint32_t)blocks[i];
        v.view = i;
        total_blocks += v.blocks;
        table.push_back(v);
        entry_view.push_back(i);
    }
}

// One batch pass over `frames` (one per view).
int batch_pass(sg_batch *b, const sg_frame *frames) {
    const uint32_t n = (uint32_t)b->views.size();
    std::vector<PassPlan> plans(n);
    for (uint32_t i = 0; i < n; ++i) {
        sg_ctx *ctx = b->views[i];
        join_batch(ctx);
        int rc = prepare_pass(ctx, &frames[i], plans[i]);
        if (rc != SG_OK) return bfail(b, rc, "view " + std::to_string(i) + ": " + ctx->error);
    }
    std::vector<BatchView> table;
    std::vector<uint32_t> entry_view;
    uint32_t total_blocks = 0;
    build_table(b, plans, table, entry_view, total_blocks);
    // 32-bit block indices; each view is below 2^32 pixels by its size types.
    uint64_t check = 0;
    for (const BatchView &v : table) check += v.blocks;
    if (check >= (1ull << 31)) return bfail(b, SG_E_CAPACITY, "batch pass exceeds 2^31 pixel blocks");
    for (uint32_t i = 0; i < n; ++i) {
        if (!plans[i].queued) continue;
        sg_ctx *ctx = b->views[i];
        SG_BHIP(b, hipEventRecord(ctx->ev_ready, ctx->stream));
        SG_BHIP(b, hipStreamWaitEvent(b->stream, ctx->ev_ready, 0));
    }
    if (!table.empty()) {
        const bool same = b->table_valid && table.size() == b->expected.size() &&
                          std::memcmp(table.data(), b->expected.data(), table.size() * sizeof(BatchView)) == 0;
        if (!same) {
            // The staging buffer may still feed the previous pass's copy.
            SG_BHIP(b, hipStreamSynchronize(b->stream));
            std::memcpy(b->staging, table.data(), table.size() * sizeof(BatchView));
            SG_BHIP(b, hipMemcpyAsync(b->table, b->staging, table.size() * sizeof(BatchView), hipMemcpyHostToDevice,
                                      b->stream));
            if (table.size() != b->entry_view.size() || total_blocks != b->total_blocks) b->graph_valid = false;
            b->entry_view = entry_view;
            b->total_blocks = total_blocks;
            b->table_valid = true;
        }
        if (b->profiling) SG_BHIP(b, hipEventRecord(b->ev_start, b->stream));
        if (b->launch.use_graph) {
            if (!b->graph_valid) {
                int rc = capture_pass(b);
                if (rc != SG_OK) return rc;
            }
            SG_BHIP(b, hipGraphLaunch(b->exec, b->stream));
        } else {
            int rc = queue_pass(b);
            if (rc != SG_OK) return rc;
        }
        if (b->profiling) {
            SG_BHIP(b, hipEventRecord(b->ev_stop, b->stream));
            SG_BHIP(b, hipEventSynchronize(b->ev_stop));
            float ms = 0.f;
            SG_BHIP(b, hipEventElapsedTime(&ms, b->ev_start, b->ev_stop));
            b->kernel_ms += ms;
            b->launches += 1;
        }
        for (BatchView &v : table) {
            v.ta.accumulated += v.ta.count;
            v.ta.sequence += v.ta.count;
            v.scene.s.counts[1] += v.ta.count;
            v.scene.s.sampling[0] += v.ta.count;
        }
        b->expected.swap(table);
        b->pass_queued = true;
        SG_BHIP(b, hipEventRecord(b->ev_done, b->stream));
        for (uint32_t i : b->entry_view) b->views[i]->batch_event = b->ev_done;
    }
    for (uint32_t i = 0; i < n; ++i) finish_pass(b->views[i], plans[i].count);
    return SG_OK;
}

// Forget the views' progressive history so the next pass starts afresh,
// with each view's noise sequence back at `sequences`.
void reset_history(sg_batch *b, const std::vector<uint32_t> &sequences) {
    for (size_t i = 0; i < b->views.size(); ++i) {
        sg_ctx *ctx = b->views[i];
        ctx->key.clear();
        ctx->samples = 0;
        ctx->sequence = sequences[i];
        ctx->last.count = 0;
        ctx->frame_on_device = false;
    }
    b->table_valid = false;
}

}  // namespace

extern "C" {

int sg_batch_create(sg_batch **out, sg_ctx *const *views, uint32_t count) {
    if (!out) return SG_E_INVALID;
    *out = nullptr;
    if (!views || count == 0 || count > kMaxBatchViews) return SG_E_INVALID;
    for (uint32_t i = 0; i < count; ++i) {
        if (!views[i] || views[i]->batch) return SG_E_INVALID;
        for (uint32_t j = 0; j < i; ++j)
            if (views[j] == views[i]) return SG_E_INVALID;
    }
    sg_batch *b = new sg_batch();
    auto bail = [&](int code) {
        sg_batch_destroy(b);
        return code;
    };
    if (hipGetDevice(&b->device) != hipSuccess) return bail(SG_E_DEVICE);
    for (uint32_t i = 0; i < count; ++i)
        if (views[i]->device != b->device) return bail(SG_E_INVALID);
    int cus = 0, khz = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) != hipSuccess || cus <= 0 || khz <= 0)
        return bail(SG_E_DEVICE);
    b->compute_units = (uint32_t)cus;
    b->clock_khz = (double)khz;
    b->launch = sg_batch_launch{256u, 2u * (uint32_t)cus, 1u, SG_BATCH_ORDER_GIVEN, 1u};
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->ev_start) != hipSuccess || hipEventCreate(&b->ev_stop) != hipSuccess)
        return bail(SG_E_DEVICE);
    if (hipMalloc(&b->table, kMaxBatchViews * sizeof(BatchView)) != hipSuccess ||
        hipMalloc(&b->counter, (kMaxBatchViews + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->done, (kMaxBatchViews + 2) * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&b->staging, kMaxBatchViews * sizeof(BatchView), hipHostMallocDefault) != hipSuccess)
        return bail(SG_E_DEVICE);
    b->views.assign(views, views + count);
    for (sg_ctx *ctx : b->views) ctx->batch = b;
    *out = b;
    return SG_OK;
}

int sg_batch_destroy(sg_batch *b) {
    if (!b) return SG_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (sg_ctx *ctx : b->views) {
        ctx->batch = nullptr;
        if (ctx->batch_event == b->ev_done) ctx->batch_event = nullptr;
    }
    drop_graph(b);
    for (hipEvent_t e : b->view_events) (void)hipEventDestroy(e);
    if (b->table) (void)hipFree(b->table);
    if (b->counter) (void)hipFree(b->counter);
    if (b->done) (void)hipFree(b->done);
    if (b->staging) (void)hipHostFree(b->staging);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->ev_stop) (void)hipEventDestroy(b->ev_stop);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SG_OK;
}

int sg_batch_set_launch(sg_batch *b, const sg_batch_launch *launch) {
    if (!b || !launch) return SG_E_INVALID;
    sg_batch_launch l = *launch;
    if (l.workgroups == 0) l.workgroups = 2u * b->compute_units;
    if (!valid_launch(l)) return bfail(b, SG_E_INVALID, "invalid batch launch settings");
    if (l.block != b->launch.block || l.order != b->launch.order) b->table_valid = false;
    if (std::memcmp(&l, &b->launch, sizeof(l)) != 0) b->graph_valid = false;
    b->launch = l;
    return SG_OK;
}

int sg_batch_get_launch(const sg_batch *b, sg_batch_launch *launch) {
    if (!b || !launch) return SG_E_INVALID;
    *launch = b->launch;
    return SG_OK;
}

int sg_batch_encode(sg_batch *b, const sg_frame *frames) {
    if (!b || !frames) return SG_E_INVALID;
    return batch_pass(b, frames);
}

int sg_batch_synchronize(sg_batch *b) {
    if (!b) return SG_E_INVALID;
    SG_BHIP(b, hipStreamSynchronize(b->stream));
    return SG_OK;
}

int sg_batch_view_times(sg_batch *b, double *milliseconds, uint32_t count, double *span_ms) {
    if (!b || !milliseconds || count != b->views.size()) return SG_E_INVALID;
    SG_BHIP(b, hipStreamSynchronize(b->stream));
    for (uint32_t i = 0; i < count; ++i) milliseconds[i] = -1.0;
    if (span_ms) *span_ms = -1.0;
    if (!b->pass_queued) return SG_OK;
    const size_t entries = b->entry_view.size();
    std::vector<unsigned long long> done(entries), clock(2);
    if (entries)
        SG_BHIP(b, hipMemcpy(done.data(), b->done, entries * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    SG_BHIP(b, hipMemcpy(clock.data(), b->done + kMaxBatchViews, 2 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    for (size_t e = 0; e < entries; ++e)
        milliseconds[b->entry_view[e]] = (double)(done[e] - clock[0]) / b->clock_khz;
    if (span_ms) *span_ms = (double)(clock[1] - clock[0]) / b->clock_khz;
    return SG_OK;
}

int sg_batch_set_profiling(sg_batch *b, int32_t enable) {
    if (!b) return SG_E_INVALID;
    b->profiling = enable != 0;
    b->kernel_ms = 0.0;
    b->launches = 0;
    return SG_OK;
}

int sg_batch_kernel_time(sg_batch *b, double *milliseconds, uint64_t *launches) {
    if (!b || !milliseconds || !launches) return SG_E_INVALID;
    *milliseconds = b->kernel_ms;
    *launches = b->launches;
    return SG_OK;
}

int sg_batch_encode_views(sg_batch *b, const sg_frame *frames, double *milliseconds, int32_t serial) {
    if (!b || !frames) return SG_E_INVALID;
    const uint32_t n = (uint32_t)b->views.size();
    if (b->view_events.size() != n) {
        for (hipEvent_t e : b->view_events) (void)hipEventDestroy(e);
        b->view_events.assign(n, nullptr);
        for (uint32_t i = 0; i < n; ++i) SG_BHIP(b, hipEventCreate(&b->view_events[i]));
    }
    SG_BHIP(b, hipStreamSynchronize(b->stream));
    for (sg_ctx *ctx : b->views) {
        join_batch(ctx);
        SG_BHIP(b, hipStreamSynchronize(ctx->stream));
    }
    SG_BHIP(b, hipEventRecord(b->ev_start, b->views[0]->stream));
    for (uint32_t i = 0; i < n; ++i) {
        sg_ctx *ctx = b->views[i];
        // One queue, as one device queue submits the views in order.
        if (serial && i > 0) SG_BHIP(b, hipStreamWaitEvent(ctx->stream, b->view_events[i - 1], 0));
        int rc = sg_encode(ctx, &frames[i]);
        if (rc != SG_OK) return bfail(b, rc, "view " + std::to_string(i) + ": " + ctx->error);
        SG_BHIP(b, hipEventRecord(b->view_events[i], ctx->stream));
    }
    for (uint32_t i = 0; i < n; ++i) {
        SG_BHIP(b, hipEventSynchronize(b->view_events[i]));
        if (milliseconds) {
            float ms = 0.f;
            SG_BHIP(b, hipEventElapsedTime(&ms, b->ev_start, b->view_events[i]));
            milliseconds[i] = ms;
        }
    }
    return SG_OK;
}

int sg_batch_autotune(sg_batch *b, const sg_frame *frames, uint32_t repeats, sg_batch_launch *best,
                      double *best_ms) {
    if (!b || !frames) return SG_E_INVALID;
    repeats = std::max<uint32_t>(1u, std::min<uint32_t>(repeats, 16u));
    const sg_batch_launch keep = b->launch;
    const bool keep_profiling = b->profiling;
    std::vector<uint32_t> sequences;
    for (const sg_ctx *ctx : b->views) sequences.push_back(ctx->sequence);
    sg_batch_launch winner = keep;
    double winner_ms = 1e300;
    const uint32_t cus = b->compute_units;
    const uint32_t blocks[] = {64u, 128u, 256u};
    const uint32_t per_cu[] = {1u, 2u, 4u, 8u};
    const uint32_t pulls[] = {1u, 4u};
    int rc = SG_OK;
    for (uint32_t block : blocks)
        for (uint32_t wpc : per_cu)
            for (uint32_t pull : pulls) {
                sg_batch_launch l{block, wpc * cus, pull, keep.order, 1u};
                if (sg_batch_set_launch(b, &l) != SG_OK) continue;
                // One untimed pass captures the graph and warms the caches.
                rc = batch_pass(b, frames);
                if (rc != SG_OK) goto out;
                sg_batch_set_profiling(b, 1);
                for (uint32_t r = 0; r < repeats; ++r) {
                    rc = batch_pass(b, frames);
                    if (rc != SG_OK) goto out;
                }
                {
                    const double ms = b->kernel_ms / (double)std::max<uint64_t>(b->launches, 1);
                    if (ms < winner_ms) {
                        winner_ms = ms;
                        winner = b->launch;
                    }
                }
                sg_batch_set_profiling(b, 0);
            }
out:
    if (rc == SG_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = bfail(b, SG_E_DEVICE, "autotune sync");
    reset_history(b, sequences);
    sg_batch_set_launch(b, rc == SG_OK ? &winner : &keep);
    sg_batch_set_profiling(b, keep_profiling ? 1 : 0);
    if (rc != SG_OK) return rc;
    if (best) *best = winner;
    if (best_ms) *best_ms = winner_ms;
    return SG_OK;
}

const char *sg_batch_last_error(const sg_batch *b) { return b ? b->error.c_str() : "null batch"; }

}  // extern "C"
