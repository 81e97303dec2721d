// Batch kernels: the persistent pass at 128 and 256 lanes, and the reset and
// advance nodes of a captured pass.
#include "sg_blocks.h"

namespace sgk {

template __global__ void k_batch<128>(BatchArgs ba);
template __global__ void k_batch<256>(BatchArgs ba);

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

}  // namespace sgk
