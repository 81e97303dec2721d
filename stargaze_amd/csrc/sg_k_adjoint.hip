// The adjoint kernel (path replay) and the gradient fill.
// The adjoint rejects frames with a ground scene or with solid-body surface
// styles: compile without them.
#define SG_LANDSCAPE 0
#define SG_STYLES 0
#include "sg_blocks.h"

namespace sgk {

__global__ void k_fill(float *p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Waves per SIMD of the adjoint kernel's register budget.
#ifndef SG_ADJOINT_WAVES
#define SG_ADJOINT_WAVES 2
#endif
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(SG_ADJOINT_WAVES, 8))) void k_adjoint(
    Scene sc, AdjointArgs aa) {
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
            // The primal's returned radiance and ring-cap scale recorded by
            // the forward pass, or the primal rerun here.
            if (aa.recorded) {
                const float4 primal = aa.recorded[(size_t)pixel * ta.count + s];
                replay_path(sc, direction, rng, accumulated, w, f3{primal.x, primal.y, primal.z}, primal.w, s_grad,
                            s_density + t, blockDim.x);
            } else {
                const PathResult r = adjoint_path(sc, direction, rng, accumulated, w, s_grad, s_density + t,
                                                  blockDim.x);
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

}  // namespace sgk
