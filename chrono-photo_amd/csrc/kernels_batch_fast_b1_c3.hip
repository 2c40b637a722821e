// kernels_batch_fast_b1_c3.hip -- group kernels of the 16-bit-halves shapes for 3-channel 8-bit stacks.
#include "k_batch.h"

namespace cp {
const void* fast_group_b1_c3(int shape, bool absolute) { return fast_group_bc<1, 3>(shape, absolute); }
}  // namespace cp
