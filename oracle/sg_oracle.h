/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C CPU restatement of stargaze's WGSL path tracer
 * (src/shaders/pathtrace.wgsl, rings.wgsl, ground.wgsl, terrain.wgsl,
 * forest.wgsl) and of the host packing in src/pathtracer.rs:789-1009. It is
 * compiled twice: with REAL = float (the reference's FP32 arithmetic, also
 * used as the CPU baseline) and with REAL = double (the FP64 oracle for
 * per-primitive hit distances). Only tests, the smoke check and the bench's
 * CPU-baseline leg may load it, and only as the checker; the product never
 * links it.
 *
 * Parity status per function is recorded in the header comment of
 * sg_oracle.c.
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H
#include <stdint.h>

typedef struct orc_scene {
    const void *settings;   /* 240-byte Settings (src/pathtracer.rs:44-52) */
    const void *bodies;     /* 160-byte Body records */
    const void *objects;    /* 48-byte Primitive records */
    const float *sky_light; /* 128 x vec4 */
    float *terrain;         /* terrain floats incl. caches (written by orc_terrain_cache) */
    const void *catalogue;  /* 32-byte RayStar records */
    uint32_t catalogue_count;
    const uint32_t *sky_cells;
    const uint32_t *milky_way;
    /* Tabulated ring optical depth (512 radial bins) of one body, or null:
     * the product's extension of src/shaders/rings.wgsl:50-81, restated. */
    const float *ring_profile;
    uint32_t ring_profile_body;
    /* Air march: view cells per camera ray and density samples per light
     * path (src/shaders/pathtrace.wgsl:67-68, replaced per frame the way
     * src/render_experiments.rs:48-70 replaces them); 0 = the shader's 8 / 4.
     * mie_height > 0 gives Mie its own scale height (an extension). */
    uint32_t air_view_steps;
    uint32_t air_sun_steps;
    float mie_height;
} orc_scene;

typedef struct orc_probe_ray {
    float offset[3];
    uint32_t anchor;
    float direction[3];
    uint32_t target;
} orc_probe_ray;

typedef struct orc_probe_result {
    double distance;
    uint32_t index;
    uint32_t pad;
    double normal[3];
    double extra[3];
} orc_probe_result;

#endif
