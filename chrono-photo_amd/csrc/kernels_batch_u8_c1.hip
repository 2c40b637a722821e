// kernels_batch_u8_c1.hip -- group kernels of the packed-byte shapes for 1-channel 8-bit stacks.
#include "k_batch.h"

namespace cp {
const void* u8_group_c1(int shape, bool absolute) { return u8_group_c<1>(shape, absolute); }
}  // namespace cp
