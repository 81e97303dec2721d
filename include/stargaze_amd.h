/*
 * stargaze_amd — C ABI of the MI355X path tracer.
 *
 * Drop-in for the GPU work that stargaze's `PathTracer::encode` issues
 * (src/pathtracer.rs:789-1009): the caller keeps building the same frame
 * (`Frame{globals, bodies, ground, scene_time}`, src/renderer.rs:50-56) and
 * hands it over through these entry points instead of wgpu buffers and
 * compute dispatches. Layouts below are byte-identical to the reference's
 * `Globals` (176 B), `Sphere`/WGSL `Body` (160 B), `Primitive` (48 B),
 * `RayStar` (32 B), `Settings` (240 B) and `CameraAir` (16 B).
 *
 * Conventions:
 *  - every function returns 0 on success and a negative SG_E* code on error
 *    (the reference panics or returns anyhow::Result instead);
 *  - host inputs are copied during the call; the library owns device memory;
 *  - work is queued on one HIP stream per context; read-back calls synchronize;
 *  - one context per host thread.
 */
#ifndef STARGAZE_AMD_H
#define STARGAZE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/pathtracer.rs:21-22, src/ground.rs:8-9, src/graphics.rs:11 */
#define SG_BODY_CAPACITY 256u
#define SG_MAX_BATCH_SAMPLES 64u
#define SG_GROUND_CAPACITY 256u
#define SG_SKY_TEXELS 128u
#define SG_MAX_SAMPLES (1u << 24)
/* Finest terrain grid this build traverses (the reference stops at 2048). */
#define SG_MAX_TERRAIN_CELLS 8192u
/* Radial bins of a tabulated ring optical-depth profile (sg_set_ring_profile). */
#define SG_RING_PROFILE_BINS 512u
/* Altitude bins of a tabulated air density profile (sg_set_density_profile). */
#define SG_DENSITY_BINS 16u

#define SG_OK 0
#define SG_E_INVALID (-1)   /* bad argument, size or layout */
#define SG_E_CAPACITY (-2)  /* too many bodies / primitives / samples */
#define SG_E_DEVICE (-3)    /* a HIP call failed; see sg_last_error */
#define SG_E_STATE (-4)     /* call made in the wrong state */

/* src/renderer.rs:13-30 `Globals`; WGSL src/shaders/pathtrace.wgsl:5-18 */
typedef struct sg_globals {
    float cam_right[4];
    float cam_up[4];
    float cam_forward[4]; /* w = tan(vertical FOV / 2) */
    float viewport[4];    /* trace width, height, exposure, display aspect */
    float atmo_center[4]; /* xyz = host centre (AU), w = host radius */
    float atmo_rayleigh[4]; /* rgb = Rayleigh (1/AU), w = Mie */
    float atmo_params[4]; /* x = Mie g, y = scale height, z = top altitude, w = host index + 1 */
    float ground_east[4]; /* w = observer height (m) */
    float ground_up[4];   /* w = host radius (m) */
    float ground_north[4]; /* w = host index + 1 (0 disables the landscape) */
    uint32_t ground_counts[4]; /* props, terrain on, theme | log2(cells) << 8, prop top bits */
} sg_globals;

/* src/renderer.rs:32-48 `Sphere`; WGSL `Body` src/shaders/pathtrace.wgsl:26-37 */
typedef struct sg_sphere {
    float center[3];
    float radius;
    float color[3];
    float emissive;
    float center_low[4]; /* xyz = residual centre, w = procedural albedo strength */
    float ring_plane[4]; /* xyz = pole, w = inner radius / body radius */
    float ring_params[4]; /* outer radius / body radius, optical depth, banded, unused */
    float ring_color[4];
    float axis_x[4];     /* w = surface style */
    float axis_y[4];     /* w = pattern seed */
    float surface_color[4];
    float surface_params[4];
} sg_sphere;

/* src/ground.rs:164-170 `Primitive`; WGSL src/shaders/ground.wgsl:3-7 */
typedef struct sg_primitive {
    float center[4]; /* w: 0 box, 1 ellipsoid, 2 rock, 2.25 tumbled, 3 conifer, 4 grass, 5 log, <0 group */
    float extent[4];
    float albedo[4];
} sg_primitive;

/* src/stars.rs:143-148 `RayStar`; WGSL `CatalogueStar` src/shaders/pathtrace.wgsl:38-41 */
typedef struct sg_ray_star {
    float direction[4]; /* xyz = unit direction, w = sin(angular radius)^2 */
    float radiance[4];
} sg_ray_star;

/* src/pathtracer.rs:44-52 `Settings`; WGSL src/shaders/pathtrace.wgsl:19-25 */
typedef struct sg_settings {
    sg_globals globals;
    uint32_t counts[4];   /* bodies, accumulated samples, samples this dispatch, max surface vertices */
    uint32_t sampling[4]; /* noise sequence */
    float background[4];  /* Milky Way brightness, map width, map height, point-source gain */
    uint32_t lights[4];   /* first three emissive indices, total count */
} sg_settings;

/* WGSL `CameraAir` src/shaders/pathtrace.wgsl:54 */
typedef struct sg_camera_air {
    float radiance[3];
    uint32_t rng;
} sg_camera_air;

/* src/pathtracer.rs:76-112 `Options` (display-only fields kept for parity). */
typedef struct sg_options {
    uint32_t noise_seed;
    uint32_t samples_per_frame;
    uint32_t max_bounces;
    uint32_t sample_limit;
    float exposure;
    int32_t auto_exposure;
    float ev_bias;
    float auto_key;
    float auto_percentile;
    float milky_way;
} sg_options;

/* src/pathtracer.rs:54-66 `Exposure`; WGSL `Exposure` src/shaders/exposure.wgsl:29-35 */
typedef struct sg_exposure {
    float value;      /* eased automatic exposure (updated on the device) */
    float key;        /* target luminance of the metered percentile */
    float percentile; /* which luminance percentile to meter (0..1) */
    float enabled;    /* 1 while automatic exposure is on */
    float bias;       /* exposure compensation in stops */
} sg_exposure;

/* src/renderer.rs:50-56 `Frame` (labels are UI-only and not taken). */
typedef struct sg_frame {
    const sg_globals *globals;
    const sg_sphere *bodies;
    uint32_t body_count;
    const sg_primitive *ground;
    uint32_t ground_count;
    double scene_time;
} sg_frame;

/* Probe rays: the reference pins geometry with probe shaders that call the
 * production helpers (src/ring_tests.rs:67-101, src/ground_tests.rs:694-726).
 * These entry points run the same device functions the integrator uses. */
typedef struct sg_probe_ray {
    float offset[3];
    uint32_t anchor; /* 0xffffffff = camera origin */
    float direction[3];
    uint32_t target; /* body index, light index, or unused */
} sg_probe_ray;

typedef struct sg_probe_result {
    float distance;
    uint32_t index;
    float normal[3];
    float extra[3];
} sg_probe_result;

enum sg_probe_kind {
    SG_PROBE_SPHERE = 0,     /* sphere_hit(ray, target)              pathtrace.wgsl:130 */
    SG_PROBE_CLOSEST = 1,    /* closest_hit(ray)                     pathtrace.wgsl:173 */
    SG_PROBE_RING = 2,       /* ring_hit + ring_transmission         rings.wgsl:8,83 */
    SG_PROBE_VISIBILITY = 3, /* light_visibility(ray, target)        rings.wgsl:90 */
    SG_PROBE_TERRAIN = 4,    /* terrain_query, metre frame           terrain.wgsl:152 */
    SG_PROBE_LANDSCAPE = 5,  /* landscape_hit(ray)                   ground.wgsl:123 */
    SG_PROBE_CONE = 6,       /* cone_width + sample_light draw       pathtrace.wgsl:197,214 */
    SG_PROBE_ALBEDO = 7,     /* style albedo of bodies[target] at the local unit point
                                q = direction, unclamped (plain style: surface_albedo),
                                the material probe of surface_tests.rs:106-125 */
    SG_PROBE_GROUND = 8,     /* ground material of landscape element `anchor` at the
                                surface-frame point `offset` (metres) with surface-frame
                                geometric normal `direction`, seen from the distance whose
                                float bits are `target`: normal = shading normal (world
                                frame), extra = albedo       ground.wgsl:597-648 */
    SG_PROBE_NOISE = 9,      /* at the point q = direction: distance = vnoise(q),
                                normal = {hash31(q), fbm(q, target)}  pathtrace.wgsl:338-365 */
    SG_PROBE_SKY = 10        /* normal = surface_sky(ray, direction, rng = hash(target + 817))
                                over the sky cache of the last encode; extra = that cache's
                                texel target % 128 (16 x 8, row-major)   forest.wgsl:3-54 */
};

typedef struct sg_ctx sg_ctx;

/* PathTracer::new (src/pathtracer.rs:292). `catalogue`/`sky_cells` are the
 * output of stars::ray_catalogue (src/stars.rs:152-205); either may be empty. */
int sg_create(sg_ctx **out, uint32_t width, uint32_t height,
              const sg_ray_star *catalogue, uint32_t catalogue_count,
              const uint32_t *sky_cells, uint32_t sky_cell_count,
              const sg_options *options);
/* Drop (src/pathtracer.rs PathTracer has no explicit destroy). */
int sg_destroy(sg_ctx *ctx);
/* Milky Way RGB10 map (src/sky.rs:103-117; bound at binding 6). */
int sg_set_milky_way(sg_ctx *ctx, const uint32_t *texels, uint32_t width, uint32_t height);
/* PathTracer::resize (src/pathtracer.rs:673) */
int sg_resize(sg_ctx *ctx, uint32_t width, uint32_t height);
/* PathTracer::samples / size / needs_redraw (src/pathtracer.rs:742-752) */
uint32_t sg_samples(const sg_ctx *ctx);
int sg_size(const sg_ctx *ctx, uint32_t *width, uint32_t *height);
int sg_needs_redraw(const sg_ctx *ctx);
/* PathTracer::prepare_frame (src/pathtracer.rs:756): 1 = fresh history. */
int sg_prepare_frame(sg_ctx *ctx, const sg_frame *frame);
/* PathTracer::set_quality (src/pathtracer.rs:761) */
int sg_set_quality(sg_ctx *ctx, uint32_t samples, uint32_t bounces, uint32_t limit);
/* PathTracer::set_exposure_controls (src/pathtracer.rs:777) */
int sg_set_exposure_controls(sg_ctx *ctx, int32_t auto_exposure, float bias);
/* Terrain::data for the ground's theme (src/terrain.rs:47-52, uploaded at
 * src/pathtracer.rs:914-932): heights, derivatives (cells <= 1024), min/max. */
int sg_upload_terrain(sg_ctx *ctx, const float *data, uint64_t float_count, uint32_t cells);
/* PathTracer::encode (src/pathtracer.rs:789): queue one progressive pass. */
int sg_encode(sg_ctx *ctx, const sg_frame *frame);
/* Display-only exposure (src/pathtracer.rs:883-906,981-1001). Every sg_encode
 * publishes the exposure controls; with auto exposure it then meters the
 * central rectangle into a 256-bin luminance histogram and eases `value`
 * towards key / percentile (src/shaders/exposure.wgsl). Batched passes do not
 * meter. sg_read_exposure returns the state and, optionally, the histogram
 * (256 counts) of the last metering. */
int sg_read_exposure(sg_ctx *ctx, sg_exposure *state, uint32_t *histogram);
/* PathTracer::display (src/pathtracer.rs:1010, src/shaders/display.wgsl:62-97):
 * the exposed, ACES-mapped image at out_width x out_height, bilinear from the
 * trace resolution, as linear display values (rgba f32, a = 1) or through the
 * sRGB transfer function as 8-bit rgba. Synchronizes the context stream. */
int sg_display(sg_ctx *ctx, uint32_t out_width, uint32_t out_height, float *rgba);
int sg_display_srgb8(sg_ctx *ctx, uint32_t out_width, uint32_t out_height, uint8_t *rgba);
/* sg_display_srgb8 without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The 8-bit target is
 * staged in a static per-GPU buffer of at most 3840 x 2160 pixels (31.6 MiB
 * on every GPU the library runs on, held until the process ends); larger
 * targets return SG_E_CAPACITY. Every pixel of the target is written, so
 * nothing carries over from the previous call. Calls that overlap on one
 * GPU (contexts of that GPU used from several threads at once) race on the
 * shared staging. */
int sg_display_srgb8_noalloc(sg_ctx *ctx, uint32_t out_width, uint32_t out_height, uint8_t *rgba);
/* Read-back of `accumulation` / `backdrop_accumulation` (vec4 f32 per pixel,
 * src/pathtracer_tests.rs:160-182). Synchronizes the context stream. */
int sg_read_accumulation(sg_ctx *ctx, float *rgba, uint64_t pixel_count);
int sg_read_backdrop(sg_ctx *ctx, float *rgba, uint64_t pixel_count);
/* Restore the images (for example from a checkpoint of a progressive
 * render); the history is unchanged. No reference counterpart. */
int sg_write_accumulation(sg_ctx *ctx, const float *rgba, uint64_t pixel_count);
int sg_write_backdrop(sg_ctx *ctx, const float *rgba, uint64_t pixel_count);
/* Per-pixel running mean of squared radiance (rgb, sample count in w); only
 * kept after sg_enable_moments(ctx, 1). The reference keeps means only; the
 * second moment is what tile-wise Monte Carlo error bars need. */
int sg_enable_moments(sg_ctx *ctx, int32_t enable);
int sg_read_moments(sg_ctx *ctx, float *rgba, uint64_t pixel_count);
int sg_synchronize(sg_ctx *ctx);
/* Device time of the integrator kernels since the last reset (HIP events on
 * the context stream), for measurement. */
int sg_set_profiling(sg_ctx *ctx, int32_t enable);
int sg_kernel_time(sg_ctx *ctx, double *milliseconds, uint64_t *launches);
/* Device time of one step of the wavefront integrator since sg_set_profiling
 * (HIP events around each launch; no reference counterpart): the extend,
 * camera-air and shade kernels, the deferred camera-air shadow rays
 * (k_air_occlusion and k_air_resolve), and how many launches the time
 * covers. */
#define SG_STEP_EXTEND 0u
#define SG_STEP_CAMERA_AIR 1u
#define SG_STEP_SHADE 2u
#define SG_STEP_AIR_OCCLUSION 3u
#define SG_STEP_COUNT 4u
int sg_step_time(const sg_ctx *ctx, uint32_t step, double *milliseconds, uint64_t *launches);
/* Integrator selection (no reference counterpart: the reference always runs
 * its megakernel). SG_INTEGRATOR_WAVEFRONT runs extend / camera-air / shade
 * kernels over compacted queues of at most `chunk_paths` live paths;
 * SG_INTEGRATOR_FUSED runs one lane per path through the whole loop;
 * SG_INTEGRATOR_AUTO (default) picks per pass: the wavefront for a frame with
 * a ground scene and air once the pass traces at least 2^20 paths, the fused
 * kernel otherwise. Both draw the same samples; images agree to the last
 * bits, and on ground scenes with air to 1e-5 (deferred camera-air shadow
 * rays are summed in their own order). `chunk_paths` 0 keeps the current value. */
#define SG_INTEGRATOR_AUTO 0u
#define SG_INTEGRATOR_FUSED 1u
#define SG_INTEGRATOR_WAVEFRONT 2u
int sg_set_integrator(sg_ctx *ctx, uint32_t integrator, uint32_t chunk_paths);
/* Compaction threshold of the wavefront integrator: once fewer than
 * `tail_paths` paths of a chunk are alive, one launch runs each of them
 * through the rest of its loop (same draws) instead of further
 * extend / shade rounds. 0 (default) compacts to the end. */
int sg_set_tail(sg_ctx *ctx, uint32_t tail_paths);
typedef struct sg_launch {
    uint32_t integrator;  /* SG_INTEGRATOR_* */
    uint32_t chunk_paths; /* wavefront queue capacity (paths in flight) */
    uint32_t tail_paths;  /* compaction threshold, see sg_set_tail */
} sg_launch;
int sg_get_launch(const sg_ctx *ctx, sg_launch *launch);
/* Records per path (1-256, default 8; 48 B each) of the wavefront
 * integrator's queue of camera-air shadow rays whose landscape query is
 * deferred to a kernel of its own (frames with a ground scene and air). A
 * chunk that pushes more grows the queue and runs again with the same
 * draws, so the value sizes memory only; sg_get_air_queue reads the current
 * size. No counterpart in the reference, whose camera-air pass
 * (atmosphere, src/shaders/pathtrace.wgsl:1092-1215) answers each query inline. */
int sg_set_air_queue(sg_ctx *ctx, uint32_t records_per_path);
int sg_get_air_queue(const sg_ctx *ctx, uint32_t *records_per_path);
/* Time trial passes of `frame` over queue capacities, compaction thresholds
 * and the fused integrator, keep the fastest and return it. Trial passes
 * render into the image; the accumulation restarts with the next pass, at
 * the noise sequence it would have had without them. */
int sg_autotune(sg_ctx *ctx, const sg_frame *frame, uint32_t repeats, sg_launch *best, double *best_ms);
const char *sg_last_error(const sg_ctx *ctx);
/* ------------------------------------------------------------------------
 * Differentiable rendering (no reference counterpart).
 *
 * sg_set_ring_profile replaces the banded optical-depth profile of one
 * body's ring (src/shaders/rings.wgsl:50-81) by SG_RING_PROFILE_BINS values
 * over the fractional radius from the inner to the outer edge (bin =
 * floor(x * bins)); the ring colour keeps its banded tint. tau = NULL
 * restores the analytic profile. Either call restarts the accumulation.
 *
 * sg_adjoint differentiates the last pass sg_encode rendered (same frame,
 * same draws): it returns d L / d parameters for
 *     L = sum over pixels p of dot(d_image[p], pass_mean[p].rgb),
 * where pass_mean is the mean radiance of that pass's samples (the image
 * itself after a pass from a fresh history). The gradient vector has
 * SG_GRAD_COUNT entries, laid out as below; parameters are the stored values
 * (sg_sphere.color, tau bins, atmo_rayleigh rgb and w, atmo_params x,
 * density bins; the density gradient is zero unless a profile is set).
 * Scenes with a landscape, or with a lit body of a solid-body surface
 * style (2-6), are rejected. Body albedos are differentiated for surface
 * styles 0 (mottled colour) and 1 (gas giant: d albedo / d zone colour,
 * material.rgb). `pass_mean` (optional,
 * W*H*4 floats) receives the pass mean the adjoint recomputed.
 * ---------------------------------------------------------------------- */
#define SG_GRAD_ALBEDO(body, channel) (3u * (body) + (channel))
#define SG_GRAD_RING(bin) (3u * SG_BODY_CAPACITY + (bin))
#define SG_GRAD_AIR (3u * SG_BODY_CAPACITY + SG_RING_PROFILE_BINS) /* +0..2 Rayleigh rgb, +3 Mie, +4 Mie g */
#define SG_GRAD_DENSITY(bin) (SG_GRAD_AIR + 5u + (bin))
#define SG_GRAD_COUNT (SG_GRAD_AIR + 5u + SG_DENSITY_BINS)

int sg_set_ring_profile(sg_ctx *ctx, uint32_t body, const float *tau, uint32_t bins);
/* Relative air density per altitude slice (SG_DENSITY_BINS equal slices from
 * the surface to atmo_params.z) replacing exp(-h / H) in every density
 * sample (src/shaders/pathtrace.wgsl:970,1166,1230); NULL restores the
 * exponential. Restarts the accumulation. */
int sg_set_density_profile(sg_ctx *ctx, const float *density, uint32_t bins);
/* Air march and Mie scale height. The reference's shader marches
 * ATMO_VIEW_STEPS = 8 view cells per camera ray (capped by its
 * quarter-scale-height rule) and ATMO_SUN_STEPS = 4 density samples per light
 * path, with one scale height for Rayleigh and Mie
 * (src/shaders/pathtrace.wgsl:67-68,944-946); its benchmark variants replace
 * the two step constants (src/render_experiments.rs:48-70). view_steps in
 * [2, 32], light_steps in [1, 32]. mie_scale_height (AU, the unit of
 * atmo_params.y) > 0 gives the Mie density its own exp(-h / H_mie) while no
 * density profile is set; 0 keeps the single scale height. NULL restores
 * (8, 4, 0). Restarts the accumulation. */
typedef struct sg_air_model {
    uint32_t view_steps;
    uint32_t light_steps;
    float mie_scale_height;
    uint32_t reserved; /* 0 */
} sg_air_model;
#define SG_AIR_MAX_STEPS 32u
int sg_set_air_model(sg_ctx *ctx, const sg_air_model *model);
int sg_get_air_model(const sg_ctx *ctx, sg_air_model *model);
/* Restart the accumulation with the next pass while keeping the noise
 * sequence running, as a changed frame does (src/pathtracer.rs:172-188):
 * every restarted pass draws fresh samples. */
int sg_restart(sg_ctx *ctx);
/* Identity of the last pass that rendered samples: it changes with every
 * such pass and with every call that restarts the accumulation (profiles,
 * air model, terrain, sg_restart). A caller that differentiates a pass later
 * records it after sg_encode and compares before sg_adjoint. */
uint64_t sg_pass_id(const sg_ctx *ctx);
int sg_adjoint(sg_ctx *ctx, const sg_frame *frame, const float *d_image, uint64_t pixel_count,
               float *gradient, uint32_t gradient_count, float *pass_mean);
/* With enable = 1, every pass also keeps each path sample's returned
 * radiance and ring-cap scale (16 B per sample), and sg_adjoint of that pass
 * replays each path once instead of rerunning its primal first. */
int sg_enable_adjoint(sg_ctx *ctx, int32_t enable);
/* Device time of sg_adjoint kernels since sg_set_profiling (HIP events). */
int sg_adjoint_time(sg_ctx *ctx, double *milliseconds, uint64_t *launches);

/* Run production device helpers on a batch of probe rays (see sg_probe_kind). */
int sg_probe(sg_ctx *ctx, const sg_frame *frame, uint32_t kind,
             const sg_probe_ray *rays, uint32_t count, sg_probe_result *results);

/* ------------------------------------------------------------------------
 * Batched views (no reference counterpart: the reference renders one view
 * per PathTracer, src/pathtracer.rs:190). A batch borrows existing contexts
 * as views of uneven resolution and sample count and renders one progressive
 * pass of all of them per sg_batch_encode, as a single persistent launch
 * replayed from a hipGraph while the set of views and their sample counts
 * stay the same. Each view keeps its own history, options and images: read
 * them with sg_read_accumulation on the view context, as after sg_encode.
 * A context belongs to at most one batch and cannot be destroyed before it.
 * ---------------------------------------------------------------------- */
typedef struct sg_batch sg_batch;

#define SG_BATCH_ORDER_GIVEN 0u    /* views' blocks in the order given */
#define SG_BATCH_ORDER_SHORTEST 1u /* fewest pixel blocks first (lowest mean latency) */
#define SG_BATCH_ORDER_LONGEST 2u  /* most pixel blocks first */

typedef struct sg_batch_launch {
    uint32_t block;           /* lanes per workgroup: 64, 128 or 256 */
    uint32_t workgroups;      /* persistent grid; 0 = two per compute unit */
    uint32_t blocks_per_pull; /* pixel blocks claimed per atomic, 1..64 */
    uint32_t order;           /* SG_BATCH_ORDER_* */
    uint32_t use_graph;       /* 1 = replay each pass from a hipGraph, 0 = plain launches */
} sg_batch_launch;

int sg_batch_create(sg_batch **out, sg_ctx *const *views, uint32_t count);
int sg_batch_destroy(sg_batch *batch);
int sg_batch_set_launch(sg_batch *batch, const sg_batch_launch *launch);
int sg_batch_get_launch(const sg_batch *batch, sg_batch_launch *launch);
/* One progressive pass of every view; frames[i] is view i's frame. Same
 * history, sample-limit and Settings semantics per view as sg_encode. */
int sg_batch_encode(sg_batch *batch, const sg_frame *frames);
int sg_batch_synchronize(sg_batch *batch);
/* Completion time of each view in the last batch pass, in ms after the first
 * workgroup started (GPU constant-rate clock); -1 for views without samples
 * in that pass. `span_ms` (optional) receives the whole pass. */
int sg_batch_view_times(sg_batch *batch, double *milliseconds, uint32_t count, double *span_ms);
int sg_batch_set_profiling(sg_batch *batch, int32_t enable);
int sg_batch_kernel_time(sg_batch *batch, double *milliseconds, uint64_t *launches);
/* The per-view form for comparison: one sg_encode per view in order, each on
 * its view's stream (serial = 0, so views may overlap on the device) or each
 * waiting for the previous view (serial = 1, one queue in submission order);
 * milliseconds[i] (optional) = completion of view i after the first launch
 * was queued. */
int sg_batch_encode_views(sg_batch *batch, const sg_frame *frames, double *milliseconds, int32_t serial);
/* Time trial passes over launch settings (workgroup size, grid size, blocks
 * per pull) and keep the fastest. Trial passes render into the views, whose
 * progressive history restarts afterwards at the noise sequence each view
 * had before. */
int sg_batch_autotune(sg_batch *batch, const sg_frame *frames, uint32_t repeats, sg_batch_launch *best,
                      double *best_ms);
const char *sg_batch_last_error(const sg_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
