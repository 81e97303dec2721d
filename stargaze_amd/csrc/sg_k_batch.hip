// The batch kernel at 128 and 256 lanes.
#include "sg_blocks.h"

namespace sgk {

template __global__ void SG_K(k_batch)<128>(BatchArgs ba);
template __global__ void SG_K(k_batch)<256>(BatchArgs ba);

}  // namespace sgk
