// Launch interface between the host library (sg_library.hip) and the kernel
// translation units (sg_k_*.hip): argument blocks and kernel declarations.
// Each kernel is defined in one translation unit; the host launches it
// through these declarations.
#pragma once
#include <hip/hip_runtime.h>

#include "sg_integrator.h"

namespace sgk {
using namespace sg;

constexpr uint32_t kTraceBlock = 256;

// Register budgets, as waves per SIMD. At one wave a kernel has the whole
// register file, and the traversals' memory latency goes unhidden; the
// budgets below were chosen by timing on MI355X, spills included.
//   fused and batch kernels: 2 with the landscape (about 130 registers
//     spilled), 3 without it (about 80 spilled);
//   k_extend: 4 (128 registers, none spilled);
//   k_camera_air, k_shade: 2 (a few dozen spilled);
//   k_tail: 1 (it runs the last few paths of a chunk).
#ifndef SG_LANDSCAPE_WAVES
#define SG_LANDSCAPE_WAVES 2
#endif
#ifndef SG_PLAIN_WAVES
#define SG_PLAIN_WAVES 3
#endif
#if SG_LANDSCAPE
#define SG_FUSED_WAVES SG_LANDSCAPE_WAVES
#else
#define SG_FUSED_WAVES SG_PLAIN_WAVES
#endif
#ifndef SG_TAIL_WAVES
#define SG_TAIL_WAVES 1
#endif
#ifndef SG_EXTEND_WAVES
#define SG_EXTEND_WAVES 4
#endif
#ifndef SG_SHADE_WAVES
#define SG_SHADE_WAVES 2
#endif
#ifndef SG_AIR_WAVES
#define SG_AIR_WAVES 2
#endif
#define SG_WAVES(n) __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(n, 8)))

// Regions of the deferred camera-air shadow-ray queue, each with its own
// counter, so that the queue's atomics spread over as many addresses.
constexpr uint32_t kAirRegions = 64;

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
    // Camera-air shadow rays whose landscape predicates k_air_occlusion
    // answers: records (origin.xyz, light distance), (direction.xyz, next
    // record of the same path), (contribution.xyz, predicates owed) in
    // air_a/b/c; per path, the throughput the contributions are scaled by
    // and the path's last record in air_pending. A wave pushes into region
    // blockIdx.x % kAirRegions: records [j * air_capacity, (j + 1) *
    // air_capacity) counted in air_count[j]. Null when the frame has no
    // landscape or no air.
    float4 *air_a, *air_b, *air_c;
    float4 *air_pending;
    uint32_t *air_count;
    uint32_t air_capacity;
    // Per-path results.
    float4 *value;
    float4 *backdrop;
    float4 *records;   // per path sample of the pass (see TraceArgs), or null
    uint32_t width, height;
    uint32_t count, accumulated, sequence, has_air;
    uint32_t pixel_begin, paths;
};

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
    uint32_t *counter;            // [0] next block
    unsigned long long *done;     // [e] clock at which entry e finished (max over its blocks)
    unsigned long long *clock;    // [0] first workgroup start, [1] last workgroup end
};

// ---------------------------------------------------------------------------
// Adjoint of one progressive pass (no reference counterpart). Lane layout as
// k_trace; each lane takes its path sample's primal (recorded, or rerun) and replays it (replay_path),
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

// ---------------------------------------------------------------------------
// Display-only exposure and tone mapping (src/shaders/exposure.wgsl,
// src/shaders/display.wgsl).
// ---------------------------------------------------------------------------
constexpr uint32_t kHistogramBins = 256;  // exposure.wgsl:41
constexpr float kLogMin = -24.0f, kLogMax = 24.0f;
constexpr float kMeterFloor = 1.0e-3f, kMinExposure = 1.0e-8f, kMaxExposure = 1.0e8f;

struct MeterArgs {
    const float4 *accumulation;
    uint32_t width, height;        // trace size
    uint32_t origin_x, origin_y;   // metered rectangle: [origin, size - origin)
    uint32_t meter_w, meter_h;
    uint32_t *histogram;           // kHistogramBins
};

struct DisplayArgs {
    const float4 *accumulation, *backdrop;
    uint32_t width, height;        // trace size
    uint32_t out_w, out_h;         // display size
    float manual_exposure;         // viewport.z of the last pass
    float milky_way;               // background.x
    const sg_exposure *state;
    float4 *out_linear;            // or null
    uchar4 *out_srgb;              // or null
};

// Kernels. The forward kernels (sg_k_trace.hip, sg_k_batch.hip,
// sg_k_wave_a.hip, sg_k_wave_b.hip) are compiled twice: as written, and with
// SG_LANDSCAPE=0 under the suffixes _st (with the solid-body surface
// styles) and _nl (SG_STYLES=0 too), for frames without a ground scene.
#if SG_LANDSCAPE
#define SG_K(name) name
#elif SG_STYLES
#define SG_K(name) name##_st
#else
#define SG_K(name) name##_nl
#endif
#define SG_FORWARD_KERNELS(S)                                                                      \
    __global__ void k_trace##S(Scene sc, TraceArgs ta);                                            \
    template <uint32_t B> __global__ void k_batch##S(BatchArgs ba);                                \
    __global__ void k_generate##S(Scene sc, WaveArgs wa, uint32_t *queue);                          \
    __global__ void k_extend##S(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n);          \
    __global__ void k_camera_air##S(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n);      \
    __global__ void k_air_occlusion##S(Scene sc, WaveArgs wa);                                       \
    __global__ void k_air_resolve##S(WaveArgs wa, const uint32_t *queue, uint32_t n);               \
    __global__ void k_shade##S(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n, uint32_t *next, \
                               uint32_t *next_count);                                              \
    __global__ void k_tail##S(Scene sc, WaveArgs wa, const uint32_t *queue, uint32_t n);            \
    __global__ void k_accumulate##S(WaveArgs wa, float4 *accumulation, float4 *backdrop, float4 *moments, \
                                    uint32_t pixels);
SG_FORWARD_KERNELS()
SG_FORWARD_KERNELS(_nl)
SG_FORWARD_KERNELS(_st)
// Compiled once (sg_k_adjoint.hip without the landscape, sg_k_misc.hip with it).
__global__ void k_batch_reset(BatchArgs ba);
__global__ void k_batch_advance(BatchArgs ba);
__global__ void k_fill(float *p, uint32_t n, float v);
__global__ void k_adjoint(Scene sc, AdjointArgs aa);
__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v);
__global__ void k_meter(MeterArgs ma);
__global__ void k_resolve(const uint32_t *histogram, sg_exposure *state);
__global__ void k_display(DisplayArgs da);
__global__ void k_terrain_cache(Scene sc, float *terrain);
__global__ void k_sky_cache(Scene sc, float4 *sky);
__global__ void k_probe(Scene sc, uint32_t kind, const sg_probe_ray *rays, uint32_t count, sg_probe_result *out);

}  // namespace sgk
