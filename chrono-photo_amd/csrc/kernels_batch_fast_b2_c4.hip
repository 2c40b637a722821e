// kernels_batch_fast_b2_c4.hip -- group kernels of the 16-bit-halves shapes for 4-channel 16-bit stacks.
#include "k_batch.h"

namespace cp {
const void* fast_group_b2_c4(int shape, bool absolute) { return fast_group_bc<2, 4>(shape, absolute); }
}  // namespace cp
