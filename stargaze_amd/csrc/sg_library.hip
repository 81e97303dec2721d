// Host library and C ABI of the MI355X path tracer (include/stargaze_amd.h).
//
// Launch structure (per progressive pass, one HIP stream per context); the
// kernels live in sg_k_*.hip and are declared in sg_launch.h:
//   k_terrain_cache   once per terrain upload / host radius change
//                     (cache_terrain_geometry, src/shaders/terrain.wgsl:242-285)
//   k_sky_cache       on a fresh accumulation (cache_surface_sky, forest.wgsl:3-17)
//   k_trace           the fused integrator: W*H*count path samples, one lane
//                     per (pixel, sample); lanes of one pixel share a
//                     wavefront, and the per-pixel mean and running-average
//                     blend of src/shaders/pathtrace.wgsl:1645-1675 are formed
//                     in LDS in the reference's summation order
//   k_generate .. k_accumulate   the wavefront integrator (default)
//   k_meter, k_resolve           automatic exposure, when enabled
// Batched views run k_batch; the adjoint runs k_adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sg_bvh.h"
#include "sg_launch.h"

using namespace sg;
using namespace sgk;

// Lanes per workgroup of the forward fused kernel (at most kTraceBlock).
constexpr uint32_t kFusedBlock = 64u;

namespace {
// The forward kernels for frames with a ground scene, and the ones compiled
// without the landscape for every other frame (same results, fewer live
// registers), with or without the solid-body surface styles (2-6).
struct ForwardKernels {
    decltype(&k_trace) trace;
    decltype(&k_batch<64>) batch64, batch128, batch256;
    decltype(&k_generate) generate;
    decltype(&k_extend) extend;
    decltype(&k_camera_air) camera_air;
    decltype(&k_air_occlusion) air_occlusion;
    decltype(&k_air_resolve) air_resolve;
    decltype(&k_shade) shade;
    decltype(&k_tail) tail;
    decltype(&k_accumulate) accumulate;
};
const ForwardKernels kLandscapeKernels{k_trace, k_batch<64>, k_batch<128>, k_batch<256>, k_generate, k_extend,
                                       k_camera_air, k_air_occlusion, k_air_resolve, k_shade, k_tail,
                                       k_accumulate};
const ForwardKernels kPlainKernels{k_trace_nl, k_batch_nl<64>, k_batch_nl<128>, k_batch_nl<256>, k_generate_nl,
                                   k_extend_nl, k_camera_air_nl, k_air_occlusion_nl, k_air_resolve_nl, k_shade_nl,
                                   k_tail_nl, k_accumulate_nl};
const ForwardKernels kStyledKernels{k_trace_st, k_batch_st<64>, k_batch_st<128>, k_batch_st<256>, k_generate_st,
                                    k_extend_st, k_camera_air_st, k_air_occlusion_st, k_air_resolve_st, k_shade_st,
                                    k_tail_st, k_accumulate_st};
const ForwardKernels &forward_kernels(bool landscape, bool styled) {
    return landscape ? kLandscapeKernels : (styled ? kStyledKernels : kPlainKernels);
}
const ForwardKernels &forward_kernels(const Scene &sc) {
    return forward_kernels(sc.s.globals.ground_north[3] != 0.0f, sc.solid_styles != 0u);
}
// 1 when a lit body of the frame has a solid-body style (2-6: cratered,
// rocky, icy, volcanic, terran).
uint32_t solid_styles(const sg_frame *frame) {
    for (uint32_t i = 0; i < frame->body_count; ++i) {
        const uint32_t style = (uint32_t)(frame->bodies[i].axis_x[3] + 0.5f);
        if (frame->bodies[i].emissive < 0.5f && style >= 2u) return 1u;
    }
    return 0u;
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
    Centre *centres = nullptr;           // FP64 body centres (device), from staged_centres
    std::vector<Centre> staged_centres;  // host copy of the last upload
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
    uint32_t integrator = SG_INTEGRATOR_AUTO;
    uint32_t chunk_paths = 1u << 23;
    uint32_t tail_paths = 0;  // compaction threshold: fused tail below this many live paths
    uint32_t wave_capacity = 0;
    void *wave_block = nullptr;
    WaveArgs wave{};
    uint32_t *queues[2] = {nullptr, nullptr};
    uint32_t *counter = nullptr;
    uint32_t *counter_host = nullptr;
    // Deferred landscape queries of camera-air shadow rays (frames with a
    // ground scene and air): three float4 per record, a pending float4 per
    // path, air_per_path records per path of the chunk.
    void *air_block = nullptr;
    size_t air_records = 0;
    uint32_t air_paths = 0;
    uint32_t air_per_path = 8;
    // Measurement
    bool profiling = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    double kernel_ms = 0.0;
    uint64_t launches = 0;
    // Per-step device time of the wavefront kernels (SG_STEP_*), from event
    // pairs around each launch while profiling.
    hipEvent_t ev_step[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double step_ms[SG_STEP_COUNT] = {};
    uint64_t step_launches[SG_STEP_COUNT] = {};
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
    // Air march (view cells, light-path samples) and Mie scale height.
    uint32_t air_view_steps = ATMO_VIEW_STEPS, air_sun_steps = ATMO_SUN_STEPS;
    float mie_height = 0.0f;
    // Identity of the last pass that rendered samples (sg_pass_id): a new
    // serial for every such pass and for every restart.
    uint64_t pass_serial = 0, pass_id = 0;
    // The last queued pass, which sg_adjoint differentiates.
    struct {
        uint32_t count = 0, samples = 0, sequence = 0;
        bool recorded = false;  // its primal outputs are in `records`
    } last;
    // Per path sample primal outputs of the last pass, kept while record_on.
    bool record_on = false;
    float4 *records = nullptr;
    uint64_t record_capacity = 0;
    // Display-only exposure state and luminance histogram on the device; the
    // manual exposure and Milky Way gain of the last pass (viewport.z, background.x).
    sg_exposure *exposure = nullptr;
    uint32_t *histogram = nullptr;
    float display_exposure = 1.0f;
    float display_milky_way = 0.0f;
    // Adjoint buffers: d(loss)/d(image) rgb, gradient, recomputed pass mean.
    float *adj_image = nullptr;
    float *adj_gradient = nullptr;
    float4 *adj_primal = nullptr;
    uint64_t adj_pixels = 0;
    double adjoint_ms = 0.0;
    uint64_t adjoint_launches = 0;
};

namespace {

// Staging of sg_display_srgb8_noalloc: one 8-bit RGBA display target of at
// most kNoAllocPixels pixels, resident on every GPU the library runs on
// (31.6 MiB each) until the process ends. Host code never uses the array's
// name as an address: each device's copy is at the address hipGetSymbolAddress
// returns with that device current (noalloc_scratch).
constexpr uint64_t kNoAllocPixels = 3840ull * 2160ull;
static __device__ __align__(256) uchar4 scratch[kNoAllocPixels];
std::mutex g_scratch_mutex;
std::vector<void *> g_scratch;  // device address of `scratch`, per device ordinal

// The address of `scratch` on `device`. With `fetch` (set-up, `device`
// current) the first call on a device asks the runtime and keeps it; later
// calls, and every call without `fetch`, return the kept address.
int noalloc_scratch(int device, bool fetch, void **ptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0) return SG_E_INVALID;
    if ((size_t)device >= g_scratch.size()) {
        if (!fetch) return SG_E_STATE;
        g_scratch.resize((size_t)device + 1, nullptr);
    }
    if (!g_scratch[device]) {
        if (!fetch) return SG_E_STATE;
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SG_E_DEVICE;
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return SG_OK;
}

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

// Copy the frame's bodies to the device with their FP64 centres.
int upload_bodies(sg_ctx *ctx, const sg_frame *frame) {
    ctx->staged_centres.resize(frame->body_count);
    fill_centres(reinterpret_cast<const Body *>(frame->bodies), frame->body_count, ctx->staged_centres.data());
    SG_HIP(ctx, hipMemcpyAsync(ctx->bodies, frame->bodies, frame->body_count * sizeof(Body), hipMemcpyHostToDevice,
                               ctx->stream));
    SG_HIP(ctx, hipMemcpyAsync(ctx->centres, ctx->staged_centres.data(), frame->body_count * sizeof(Centre),
                               hipMemcpyHostToDevice, ctx->stream));
    return SG_OK;
}

Scene make_scene(const sg_ctx *ctx, const sg_settings &settings) {
    Scene sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.s = settings;
    sc.bodies = ctx->bodies;
    sc.centres = ctx->centres;
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
    sc.air_view_steps = ctx->air_view_steps;
    sc.air_sun_steps = ctx->air_sun_steps;
    sc.mie_height = ctx->mie_height;
    return sc;
}

// Restart the accumulation with the next pass (a new scene), retiring the
// identity of the last pass.
void restart_history(sg_ctx *ctx) {
    ctx->key.clear();
    ctx->samples = 0;
    ctx->last.count = 0;
    ctx->pass_id = ++ctx->pass_serial;
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
    const sg_globals *g = frame->globals;
    if (g->ground_north[3] != 0.0f) {
        const uint32_t theme = g->ground_counts[2] & 255u;
        if (theme == 4u || theme > 5u)
            return fail(ctx, SG_E_INVALID, "ground theme 4 (basalt, which needs the volcano survey) is not built");
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

// First slot of the deferred-record counts (one per queue region) in the
// counter block.
constexpr uint32_t kAirCountSlot = 64u;

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
    SG_HIP(ctx, hipMalloc(&ctx->queues[0], (2 * n + kAirCountSlot + kAirRegions) * sizeof(uint32_t)));
    ctx->queues[1] = ctx->queues[0] + n;
    ctx->counter = ctx->queues[0] + 2 * n;
    if (!ctx->counter_host)
        SG_HIP(ctx, hipHostMalloc(&ctx->counter_host, (kAirCountSlot + kAirRegions) * sizeof(uint32_t),
                                  hipHostMallocDefault));
    ctx->wave_capacity = paths;
    return SG_OK;
}


// Record storage for `paths` paths at ctx->air_per_path records each. A path
// at the ground of the 4K Halo frame averages about two; a chunk that pushes
// more grows the storage and runs again (encode_wavefront).
int ensure_air(sg_ctx *ctx, uint32_t paths) {
    const size_t per_region = ((size_t)paths * ctx->air_per_path + kAirRegions - 1) / kAirRegions;
    const size_t records = per_region * kAirRegions;
    if (records > 0xFFFFFFF0ull)  // record indices and the list's end mark are 32-bit
        return fail(ctx, SG_E_CAPACITY, "camera-air shadow-ray queue over 2^32 records; lower the queue size");
    if (paths <= ctx->air_paths && records <= ctx->air_records && ctx->air_block) return SG_OK;
    if (ctx->air_block) SG_HIP(ctx, hipFree(ctx->air_block));
    ctx->air_block = nullptr;
    ctx->air_paths = 0;
    ctx->air_records = 0;
    SG_HIP(ctx, hipMalloc(&ctx->air_block, records * 48 + (size_t)paths * 16));
    ctx->air_paths = paths;
    ctx->air_records = records;
    return SG_OK;
}

// One progressive pass through the wavefront kernels, chunked over pixels so
// the path state stays within wave_capacity.
int encode_wavefront(sg_ctx *ctx, const Scene &sc, const TraceArgs &ta) {
    const ForwardKernels &fk = forward_kernels(sc);
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
    // Camera-air shadow rays of a ground-scene frame (the kernels compiled
    // with the landscape) queue their landscape predicates for
    // k_air_occlusion, which answers them together, coherently, instead of
    // in the divergent camera-air kernel.
    const bool defer_air = ta.has_air && &fk == &kLandscapeKernels;
    uint32_t air_records = 0;
    auto bind_air = [&]() -> int {
        if (!defer_air) {
            wa.air_a = wa.air_b = wa.air_c = wa.air_pending = nullptr;
            wa.air_count = nullptr;
            wa.air_capacity = 0;
            return SG_OK;
        }
        const int arc = ensure_air(ctx, capacity);
        if (arc != SG_OK) return arc;
        const size_t records = ctx->air_records;
        float4 *base = static_cast<float4 *>(ctx->air_block);
        wa.air_a = base;
        wa.air_b = base + records;
        wa.air_c = base + 2 * records;
        wa.air_pending = base + 3 * records;
        wa.air_count = ctx->counter + kAirCountSlot;
        air_records = (uint32_t)(records / kAirRegions);
        wa.air_capacity = air_records;
        return SG_OK;
    };
    rc = bind_air();
    if (rc != SG_OK) return rc;
    // One wave per workgroup: the path kernels share nothing across waves,
    // and a small workgroup frees its slot as soon as its one wave ends
    // (7% faster than 256 lanes on the 4K Halo pass).
    const uint32_t block = 64;
    for (uint64_t p0 = 0; p0 < pixels;) {
        const uint32_t chunk = (uint32_t)std::min<uint64_t>(chunk_pixels, pixels - p0);
        wa.pixel_begin = (uint32_t)p0;
        wa.paths = chunk * count;
        uint32_t n = wa.paths;
        int q = 0;
        bool redo = false;
        hipLaunchKernelGGL(fk.generate, dim3((n + block - 1) / block), dim3(block), 0, ctx->stream, sc, wa,
                           ctx->queues[q]);
        SG_HIP(ctx, hipGetLastError());
        for (uint32_t iteration = 0; n > 0; ++iteration) {
            if (n < ctx->tail_paths) {
                hipLaunchKernelGGL(fk.tail, dim3((n + block - 1) / block), dim3(block), 0, ctx->stream, sc, wa,
                                   ctx->queues[q], n);
                SG_HIP(ctx, hipGetLastError());
                break;
            }
            const dim3 grid((n + block - 1) / block);
            const bool air = iteration == 0 && ta.has_air;
            const bool deferred = air && defer_air;
            hipEvent_t *ev = ctx->ev_step;
            if (deferred) SG_HIP(ctx, hipMemsetAsync(wa.air_count, 0, kAirRegions * sizeof(uint32_t), ctx->stream));
            if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
            hipLaunchKernelGGL(fk.extend, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n);
            SG_HIP(ctx, hipGetLastError());
            if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
            if (air) {
                hipLaunchKernelGGL(fk.camera_air, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n);
                SG_HIP(ctx, hipGetLastError());
            }
            if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
            if (deferred) {
                // Per region, grid-stride over the records queued there (at
                // most air_records).
                const dim3 ogrid(std::min<uint32_t>((air_records + block - 1) / block, 256u), kAirRegions);
                hipLaunchKernelGGL(fk.air_occlusion, ogrid, dim3(block), 0, ctx->stream, sc, wa);
                SG_HIP(ctx, hipGetLastError());
                hipLaunchKernelGGL(fk.air_resolve, grid, dim3(block), 0, ctx->stream, wa, ctx->queues[q], n);
                SG_HIP(ctx, hipGetLastError());
            }
            SG_HIP(ctx, hipMemsetAsync(ctx->counter, 0, sizeof(uint32_t), ctx->stream));
            if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
            hipLaunchKernelGGL(fk.shade, grid, dim3(block), 0, ctx->stream, sc, wa, ctx->queues[q], n,
                               ctx->queues[q ^ 1], ctx->counter);
            SG_HIP(ctx, hipGetLastError());
            if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
            const uint32_t words = deferred ? kAirCountSlot + kAirRegions : 1;
            SG_HIP(ctx, hipMemcpyAsync(ctx->counter_host, ctx->counter, words * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, ctx->stream));
            SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            uint64_t fullest = 0;
            if (deferred)
                for (uint32_t j = 0; j < kAirRegions; ++j)
                    fullest = std::max<uint64_t>(fullest, ctx->counter_host[kAirCountSlot + j]);
            if (fullest > air_records) {
                // Pushes past a region's end were dropped: grow the queue and
                // run the chunk again. Every kernel before k_accumulate
                // rewrites what it leaves, and the pushes land in the same
                // regions, so the second run fits and its results are those
                // of a run that never overflowed.
                const uint64_t need = (fullest * kAirRegions + capacity - 1) / capacity;
                const uint64_t per_path = need + need / 4 + 1;
                if (per_path * capacity > 0xFFFFFFF0ull - kAirRegions)
                    return fail(ctx, SG_E_CAPACITY, "camera-air shadow rays exceed the deferral queue");
                ctx->air_per_path = (uint32_t)per_path;
                SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
                rc = bind_air();
                if (rc != SG_OK) return rc;
                redo = true;
                break;
            }
            if (ctx->profiling) {
                float ms[4] = {0.f, 0.f, 0.f, 0.f};
                for (int k = 0; k < 4; ++k) SG_HIP(ctx, hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
                ctx->step_ms[SG_STEP_EXTEND] += ms[0];
                ctx->step_launches[SG_STEP_EXTEND] += 1;
                if (air) {
                    ctx->step_ms[SG_STEP_CAMERA_AIR] += ms[1];
                    ctx->step_launches[SG_STEP_CAMERA_AIR] += 1;
                }
                if (deferred) {
                    ctx->step_ms[SG_STEP_AIR_OCCLUSION] += ms[2];
                    ctx->step_launches[SG_STEP_AIR_OCCLUSION] += 1;
                }
                ctx->step_ms[SG_STEP_SHADE] += ms[3];
                ctx->step_launches[SG_STEP_SHADE] += 1;
            }
            n = *ctx->counter_host;
            q ^= 1;
        }
        if (redo) continue;
        hipLaunchKernelGGL(fk.accumulate, dim3((chunk + block - 1) / block), dim3(block), 0, ctx->stream, wa,
                           ta.accumulation, ta.backdrop, ta.moments, chunk);
        SG_HIP(ctx, hipGetLastError());
        p0 += chunk_pixels;
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
        if (frame->body_count) {
            const int rc_ = upload_bodies(ctx, frame);
            if (rc_ != SG_OK) return rc_;
        }
        if (frame->ground_count)
            SG_HIP(ctx, hipMemcpyAsync(ctx->objects, frame->ground, frame->ground_count * sizeof(sg_primitive),
                                       hipMemcpyHostToDevice, ctx->stream));
        ctx->frame_on_device = true;
        plan.queued = true;
    }
    int brc = ensure_bvh(ctx, frame);
    if (brc != SG_OK) return brc;
    plan.sc = make_scene(ctx, st);
    plan.sc.solid_styles = solid_styles(frame);
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
    // One wave per workgroup for the forward pass (the samples of a pixel
    // still share it): 10% faster than 256 lanes on the ringed giant.
    ta.pixels_per_block = std::max<uint32_t>(1u, kFusedBlock / count);
    ctx->last.count = count;
    ctx->last.samples = ctx->samples;
    ctx->last.sequence = st.sampling[0];
    ctx->last.recorded = ta.records != nullptr;
    ctx->pass_id = ++ctx->pass_serial;
    return SG_OK;
}

// Publish the exposure controls (src/pathtracer.rs:883-906): while automatic
// metering stays on, only the fields after the device-eased `value`.
int publish_exposure(sg_ctx *ctx) {
    const sg_options &o = ctx->options;
    sg_exposure e{o.exposure, o.auto_key, o.auto_percentile, o.auto_exposure ? 1.0f : 0.0f, o.ev_bias};
    if (o.auto_exposure && ctx->exposure_was_enabled)
        SG_HIP(ctx, hipMemcpyAsync(&ctx->exposure->key, &e.key, 4 * sizeof(float), hipMemcpyHostToDevice,
                                   ctx->stream));
    else
        SG_HIP(ctx, hipMemcpyAsync(ctx->exposure, &e, sizeof(e), hipMemcpyHostToDevice, ctx->stream));
    ctx->display_exposure = o.exposure;
    ctx->display_milky_way = o.milky_way;
    return SG_OK;
}

// Meter the central rectangle and resolve the eased exposure
// (src/pathtracer.rs:981-1001), also once the sample limit is reached.
int meter_exposure(sg_ctx *ctx) {
    MeterArgs ma;
    ma.accumulation = ctx->accumulation;
    ma.width = ctx->width;
    ma.height = ctx->height;
    ma.origin_x = ctx->width / 4u;
    ma.origin_y = ctx->height / 4u;
    ma.meter_w = ctx->width - 2u * ma.origin_x;
    ma.meter_h = ctx->height - 2u * ma.origin_y;
    ma.histogram = ctx->histogram;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kHistogramBins), 0, ctx->stream, ctx->histogram, kHistogramBins, 0u);
    SG_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_meter, dim3((ma.meter_w + 15u) / 16u, (ma.meter_h + 15u) / 16u), dim3(16, 16), 0,
                       ctx->stream, ma);
    SG_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_resolve, dim3(1), dim3(kHistogramBins), 0, ctx->stream, ctx->histogram, ctx->exposure);
    SG_HIP(ctx, hipGetLastError());
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
    {
        // Set-up of the allocation-free display staging, once per device.
        void *staging = nullptr;
        if (noalloc_scratch(ctx->device, true, &staging) != SG_OK) return bail(SG_E_DEVICE);
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bail(SG_E_DEVICE);
    for (hipEvent_t &e : ctx->ev_step)
        if (hipEventCreate(&e) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_ready, hipEventDisableTiming) != hipSuccess)
        return bail(SG_E_DEVICE);
    if (alloc_image(ctx) != SG_OK) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->bodies, SG_BODY_CAPACITY * sizeof(Body)) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->centres, SG_BODY_CAPACITY * sizeof(Centre)) != hipSuccess) return bail(SG_E_DEVICE);
    if (ensure_objects(ctx, SG_GROUND_CAPACITY) != SG_OK) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->sky_light, SG_SKY_TEXELS * sizeof(float4)) != hipSuccess) return bail(SG_E_DEVICE);
    if (hipMalloc(&ctx->exposure, sizeof(sg_exposure)) != hipSuccess ||
        hipMalloc(&ctx->histogram, kHistogramBins * sizeof(uint32_t)) != hipSuccess)
        return bail(SG_E_DEVICE);
    {
        const sg_options &o = ctx->options;
        sg_exposure e{o.exposure, o.auto_key, o.auto_percentile, o.auto_exposure ? 1.0f : 0.0f, o.ev_bias};
        if (hipMemcpy(ctx->exposure, &e, sizeof(e), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(ctx->histogram, 0, kHistogramBins * sizeof(uint32_t)) != hipSuccess)
            return bail(SG_E_DEVICE);
        ctx->display_exposure = o.exposure;
    }
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
    void *buffers[] = {ctx->accumulation, ctx->backdrop, ctx->moments, ctx->bodies, ctx->centres, ctx->objects, ctx->sky_light,
                       ctx->terrain, ctx->catalogue, ctx->sky_cells, ctx->milky_way, ctx->probe_rays,
                       ctx->probe_results, ctx->wave_block, ctx->queues[0], ctx->air_block, ctx->bvh_nodes,
                       ctx->bvh_prims, ctx->ring_profile, ctx->density_profile, ctx->adj_image,
                       ctx->adj_gradient, ctx->adj_primal, ctx->records, ctx->exposure, ctx->histogram};
    if (ctx->counter_host) (void)hipHostFree(ctx->counter_host);
    for (void *b : buffers)
        if (b) (void)hipFree(b);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    for (hipEvent_t e : ctx->ev_step)
        if (e) (void)hipEventDestroy(e);
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
    restart_history(ctx);
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
    // New ground is a new scene: restart the accumulation, as the profile
    // setters do.
    restart_history(ctx);
    return SG_OK;
}

int sg_enable_moments(sg_ctx *ctx, int32_t enable) {
    if (!ctx) return SG_E_INVALID;
    bool on = enable != 0;
    if (on == ctx->moments_on) return SG_OK;
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->moments_on = on;
    restart_history(ctx);
    return alloc_image(ctx);
}

// The integrator a pass runs. SG_INTEGRATOR_AUTO takes the wavefront for a
// ground scene with air (its camera-air shadow rays are deferred to their
// own kernel) once the pass traces at least kAutoWavefrontPaths paths, and
// the fused kernel otherwise: measured on MI355X, the wavefront runs such
// passes of 2M-33M paths 1.1-1.5x faster, while at 0.5M paths, and for
// frames without a ground scene or without air at any size, the fused kernel
// is faster (1.2-3x).
constexpr uint64_t kAutoWavefrontPaths = 1ull << 20;
uint32_t pass_integrator(const sg_ctx *ctx, const Scene &sc, const TraceArgs &ta, uint64_t paths) {
    if (ctx->integrator != SG_INTEGRATOR_AUTO) return ctx->integrator;
    const bool deferred_air = ta.has_air && &forward_kernels(sc) == &kLandscapeKernels;
    return deferred_air && paths >= kAutoWavefrontPaths ? SG_INTEGRATOR_WAVEFRONT : SG_INTEGRATOR_FUSED;
}

// PathTracer::encode (src/pathtracer.rs:789-1009).
int sg_encode(sg_ctx *ctx, const sg_frame *frame) {
    if (!ctx) return SG_E_INVALID;
    join_batch(ctx);
    PassPlan plan;
    int rc = prepare_pass(ctx, frame, plan);
    if (rc != SG_OK) return rc;
    rc = publish_exposure(ctx);
    if (rc != SG_OK) return rc;
    if (plan.count > 0) {
        const TraceArgs &ta = plan.ta;
        const uint32_t block = ta.pixels_per_block * plan.count;
        const uint64_t pixels = (uint64_t)ctx->width * ctx->height;
        const uint32_t grid = (uint32_t)((pixels + ta.pixels_per_block - 1) / ta.pixels_per_block);
        if (ctx->profiling) SG_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
        if (pass_integrator(ctx, plan.sc, ta, pixels * plan.count) == SG_INTEGRATOR_FUSED) {
            hipLaunchKernelGGL(forward_kernels(plan.sc).trace, dim3(grid), dim3(block), 0, ctx->stream, plan.sc, ta);
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
    if (ctx->options.auto_exposure) {
        rc = meter_exposure(ctx);
        if (rc != SG_OK) return rc;
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

static int write_image(sg_ctx *ctx, float4 *dst, const float *src, uint64_t pixel_count) {
    if (!ctx || !src) return SG_E_INVALID;
    if (pixel_count != (uint64_t)ctx->width * ctx->height)
        return fail(ctx, SG_E_INVALID, "pixel count does not match the image size");
    join_batch(ctx);
    SG_HIP(ctx, hipMemcpyAsync(dst, src, pixel_count * 16u, hipMemcpyHostToDevice, ctx->stream));
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}
int sg_write_accumulation(sg_ctx *ctx, const float *rgba, uint64_t n) {
    return write_image(ctx, ctx ? ctx->accumulation : nullptr, rgba, n);
}
int sg_write_backdrop(sg_ctx *ctx, const float *rgba, uint64_t n) {
    return write_image(ctx, ctx ? ctx->backdrop : nullptr, rgba, n);
}

int sg_read_exposure(sg_ctx *ctx, sg_exposure *state, uint32_t *histogram) {
    if (!ctx || !state) return SG_E_INVALID;
    join_batch(ctx);
    SG_HIP(ctx, hipMemcpyAsync(state, ctx->exposure, sizeof(sg_exposure), hipMemcpyDeviceToHost, ctx->stream));
    if (histogram)
        SG_HIP(ctx, hipMemcpyAsync(histogram, ctx->histogram, kHistogramBins * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

// `staging` (8-bit targets only) is device memory of at least out_w * out_h
// texels the caller provides; null allocates it on the stream.
static int display(sg_ctx *ctx, uint32_t out_w, uint32_t out_h, float *linear, uint8_t *srgb,
                   uchar4 *staging = nullptr) {
    if (!ctx || out_w == 0 || out_h == 0 || (!linear && !srgb)) return SG_E_INVALID;
    if ((uint64_t)out_w * out_h > (1ull << 31)) return fail(ctx, SG_E_INVALID, "display target too large");
    join_batch(ctx);
    const size_t n = (size_t)out_w * out_h;
    void *out = staging;
    if (!staging) SG_HIP(ctx, hipMallocAsync(&out, n * (linear ? sizeof(float4) : sizeof(uchar4)), ctx->stream));
    DisplayArgs da;
    da.accumulation = ctx->accumulation;
    da.backdrop = ctx->backdrop;
    da.width = ctx->width;
    da.height = ctx->height;
    da.out_w = out_w;
    da.out_h = out_h;
    da.manual_exposure = ctx->display_exposure;
    da.milky_way = ctx->display_milky_way;
    da.state = ctx->exposure;
    da.out_linear = linear ? static_cast<float4 *>(out) : nullptr;
    da.out_srgb = linear ? nullptr : static_cast<uchar4 *>(out);
    hipLaunchKernelGGL(k_display, dim3((out_w + 15u) / 16u, (out_h + 15u) / 16u), dim3(16, 16), 0, ctx->stream, da);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(linear ? static_cast<void *>(linear) : static_cast<void *>(srgb), out,
                           n * (linear ? sizeof(float4) : sizeof(uchar4)), hipMemcpyDeviceToHost, ctx->stream);
    if (!staging) {
        hipError_t f = hipFreeAsync(out, ctx->stream);
        if (e == hipSuccess) e = f;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, SG_E_DEVICE, std::string("display: ") + hipGetErrorString(e));
    return SG_OK;
}
int sg_display(sg_ctx *ctx, uint32_t out_w, uint32_t out_h, float *rgba) { return display(ctx, out_w, out_h, rgba, nullptr); }
// sg_display_srgb8 through the per-device static staging: no allocation.
int sg_display_srgb8_noalloc(sg_ctx *ctx, uint32_t out_w, uint32_t out_h, uint8_t *rgba) {
    if (!ctx || !rgba || out_w == 0 || out_h == 0) return SG_E_INVALID;
    if ((uint64_t)out_w * out_h > kNoAllocPixels)
        return fail(ctx, SG_E_CAPACITY, "the allocation-free display target holds at most 3840 x 2160 pixels");
    void *staging = nullptr;
    if (noalloc_scratch(ctx->device, false, &staging) != SG_OK)
        return fail(ctx, SG_E_DEVICE, "device address of the display staging is unavailable");
    return display(ctx, out_w, out_h, nullptr, rgba, static_cast<uchar4 *>(staging));
}
int sg_display_srgb8(sg_ctx *ctx, uint32_t out_w, uint32_t out_h, uint8_t *rgba) {
    return display(ctx, out_w, out_h, nullptr, rgba);
}

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
    for (uint32_t k = 0; k < SG_STEP_COUNT; ++k) {
        ctx->step_ms[k] = 0.0;
        ctx->step_launches[k] = 0;
    }
    return SG_OK;
}

int sg_kernel_time(sg_ctx *ctx, double *milliseconds, uint64_t *launches) {
    if (!ctx || !milliseconds || !launches) return SG_E_INVALID;
    *milliseconds = ctx->kernel_ms;
    *launches = ctx->launches;
    return SG_OK;
}

int sg_step_time(const sg_ctx *ctx, uint32_t step, double *milliseconds, uint64_t *launches) {
    if (!ctx || !milliseconds || !launches || step >= SG_STEP_COUNT) return SG_E_INVALID;
    *milliseconds = ctx->step_ms[step];
    *launches = ctx->step_launches[step];
    return SG_OK;
}

int sg_set_integrator(sg_ctx *ctx, uint32_t integrator, uint32_t chunk_paths) {
    if (!ctx) return SG_E_INVALID;
    if (integrator != SG_INTEGRATOR_AUTO && integrator != SG_INTEGRATOR_FUSED &&
        integrator != SG_INTEGRATOR_WAVEFRONT)
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

int sg_set_air_queue(sg_ctx *ctx, uint32_t records_per_path) {
    if (!ctx) return SG_E_INVALID;
    if (records_per_path < 1u || records_per_path > 256u)
        return fail(ctx, SG_E_INVALID, "records per path must be 1-256");
    ctx->air_per_path = records_per_path;
    return SG_OK;
}

int sg_get_air_queue(const sg_ctx *ctx, uint32_t *records_per_path) {
    if (!ctx || !records_per_path) return SG_E_INVALID;
    *records_per_path = ctx->air_per_path;
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
    restart_history(ctx);
    ctx->sequence = keep_sequence;
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
    restart_history(ctx);
    ctx->exposure_idle_frames = 0;
    return SG_OK;
}

uint64_t sg_pass_id(const sg_ctx *ctx) { return ctx ? ctx->pass_id : 0u; }

int sg_set_air_model(sg_ctx *ctx, const sg_air_model *model) {
    if (!ctx) return SG_E_INVALID;
    sg_air_model m{ATMO_VIEW_STEPS, ATMO_SUN_STEPS, 0.0f, 0u};
    if (model) m = *model;
    if (m.view_steps < 2u || m.view_steps > SG_AIR_MAX_STEPS || m.light_steps < 1u || m.light_steps > SG_AIR_MAX_STEPS)
        return fail(ctx, SG_E_INVALID, "air steps must lie in [2, 32] (view) and [1, 32] (light)");
    if (!(m.mie_scale_height >= 0.0f) || !std::isfinite(m.mie_scale_height))
        return fail(ctx, SG_E_INVALID, "the Mie scale height must be finite and non-negative");
    join_batch(ctx);
    SG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->air_view_steps = m.view_steps;
    ctx->air_sun_steps = m.light_steps;
    ctx->mie_height = m.mie_scale_height;
    restart_history(ctx);
    return SG_OK;
}

int sg_get_air_model(const sg_ctx *ctx, sg_air_model *model) {
    if (!ctx || !model) return SG_E_INVALID;
    *model = sg_air_model{ctx->air_view_steps, ctx->air_sun_steps, ctx->mie_height, 0u};
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
    restart_history(ctx);
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
    restart_history(ctx);
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
    if (solid_styles(frame))
        return fail(ctx, SG_E_INVALID, "the adjoint does not differentiate solid-body surface styles (2-6)");
    if (ctx->mie_height > 0.0f && !ctx->density_on)
        return fail(ctx, SG_E_INVALID, "the adjoint does not differentiate a separate Mie scale height");
    if (ctx->density_on && (ctx->air_view_steps > COLUMN_BIN_SLOTS || ctx->air_sun_steps > COLUMN_BIN_SLOTS))
        return fail(ctx, SG_E_INVALID, "the density-profile adjoint records at most 16 samples per column");
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
        if (frame->body_count) {
            const int rc_ = upload_bodies(ctx, frame);
            if (rc_ != SG_OK) return rc_;
        }
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
    if (kind > SG_PROBE_SKY) return fail(ctx, SG_E_INVALID, "unknown probe kind");
    if (frame->body_count > SG_BODY_CAPACITY) return fail(ctx, SG_E_CAPACITY, "too many scene bodies");
    if (count == 0) return SG_OK;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t a = rays[i].anchor;
        if (kind == SG_PROBE_GROUND) {
            if (a != SG_LOCAL_BASE && a != SG_TERRAIN_ID && (a < SG_LOCAL_BASE || a - SG_LOCAL_BASE > frame->ground_count))
                return fail(ctx, SG_E_INVALID, "ground probe element out of range");
            continue;
        }
        if (a != SG_MISS && a >= frame->body_count) return fail(ctx, SG_E_INVALID, "probe anchor out of range");
        if ((kind == SG_PROBE_SPHERE || kind == SG_PROBE_RING || kind == SG_PROBE_VISIBILITY ||
             kind == SG_PROBE_ALBEDO) &&
            rays[i].target >= frame->body_count)
            return fail(ctx, SG_E_INVALID, "probe target out of range");
        if (kind == SG_PROBE_NOISE && rays[i].target > 16u) return fail(ctx, SG_E_INVALID, "fbm octaves above 16");
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
    // The viewport as an encode of this frame sees it (src/pathtracer.rs:812-817):
    // footprint-dependent ground detail and point discs read it.
    st.globals.viewport[3] = st.globals.viewport[0] / std::max(st.globals.viewport[1], 1.0f);
    st.globals.viewport[0] = (float)ctx->width;
    st.globals.viewport[1] = (float)ctx->height;
    st.globals.viewport[2] = ctx->options.exposure;
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
    if (frame->body_count) {
        const int rc_ = upload_bodies(ctx, frame);
        if (rc_ != SG_OK) return rc_;
    }
    if (frame->ground_count)
        SG_HIP(ctx, hipMemcpyAsync(ctx->objects, frame->ground, frame->ground_count * sizeof(sg_primitive),
                                   hipMemcpyHostToDevice, ctx->stream));
    int brc = ensure_bvh(ctx, frame);
    if (brc != SG_OK) return brc;
    Scene sc = make_scene(ctx, st);
    sc.solid_styles = solid_styles(frame);
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
    bool landscape = false;  // some view of the table has a ground scene
    bool styled = false;     // some view has a body with a solid-body style
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
    const ForwardKernels &fk = forward_kernels(b->landscape, b->styled);
    switch (b->launch.block) {
    case 64: hipLaunchKernelGGL(fk.batch64, grid, dim3(64), 0, b->stream, ba); break;
    case 128: hipLaunchKernelGGL(fk.batch128, grid, dim3(128), 0, b->stream, ba); break;
    default: hipLaunchKernelGGL(fk.batch256, grid, dim3(256), 0, b->stream, ba); break;
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
            bool landscape = false, styled = false;
            for (const BatchView &v : table) {
                landscape |= v.scene.s.globals.ground_north[3] != 0.0f;
                styled |= v.scene.solid_styles != 0u;
            }
            if (table.size() != b->entry_view.size() || total_blocks != b->total_blocks ||
                landscape != b->landscape || styled != b->styled)
                b->graph_valid = false;
            b->landscape = landscape;
            b->styled = styled;
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
        restart_history(ctx);
        ctx->sequence = sequences[i];
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

// Batched passes over views with a landscape are switched off: the one GPU
// run of such a pass after the air-model change ended in a device memory
// fault whose cause host runs of the same views (address and memory
// sanitizers) did not show. Views without a landscape (the batched-serving
// workload) are unaffected; sg_encode renders landscapes as before.
static int reject_landscape_views(sg_batch *b, const sg_frame *frames) {
    for (size_t i = 0; i < b->views.size(); ++i)
        if (frames[i].globals && frames[i].globals->ground_north[3] != 0.0f)
            return bfail(b, SG_E_INVALID, "view " + std::to_string(i) +
                                              ": batched passes of views with a landscape are switched off");
    return SG_OK;
}

int sg_batch_encode(sg_batch *b, const sg_frame *frames) {
    if (!b || !frames) return SG_E_INVALID;
    int rc = reject_landscape_views(b, frames);
    if (rc != SG_OK) return rc;
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
    if (int rc = reject_landscape_views(b, frames); rc != SG_OK) return rc;
    repeats = std::max<uint32_t>(1u, std::min<uint32_t>(repeats, 16u));
    const sg_batch_launch keep = b->launch;
    const bool keep_profiling = b->profiling;
    std::vector<uint32_t> sequences;
    for (const sg_ctx *ctx : b->views) sequences.push_back(ctx->sequence);
    sg_batch_launch winner = keep;
    double winner_ms = 1e300;
    const uint32_t cus = b->compute_units;
    const uint32_t blocks[] = {64u, 128u, 256u};
    // Resident waves per CU: two or three per SIMD, the fused kernel's
    // register budget. Fewer leave latency unhidden (64-lane workgroups at
    // eight per CU ran the 96-view workload 25% slower); more only queue.
    const uint32_t waves_per_cu[] = {8u, 12u};
    const uint32_t pulls[] = {1u, 4u, 16u};
    int rc = SG_OK;
    for (uint32_t block : blocks)
        for (uint32_t waves : waves_per_cu)
            for (uint32_t pull : pulls) {
                const uint32_t wpc = std::max<uint32_t>(1u, waves / (block / 64u));
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
