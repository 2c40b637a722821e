// kernels_batch_u8_c2.hip -- group kernels of the packed-byte shapes for 2-channel 8-bit stacks.
#include "k_batch.h"

namespace cp {
const void* u8_group_c2(int shape, bool absolute) { return u8_group_c<2>(shape, absolute); }
}  // namespace cp
