// The fused integrator kernel, and the batch kernel at 64 lanes.
#include "sg_blocks.h"

namespace sgk {

__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(SG_FUSED_WAVES, 8))) void SG_K(k_trace)(
    Scene sc, TraceArgs ta) {
    __shared__ float4 s_value[kTraceBlock];
    __shared__ float4 s_backdrop[kTraceBlock];
    trace_block(sc, ta, blockIdx.x, s_value, s_backdrop);
}

template __global__ void SG_K(k_batch)<64>(BatchArgs ba);

}  // namespace sgk
