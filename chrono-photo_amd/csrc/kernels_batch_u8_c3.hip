// kernels_batch_u8_c3.hip -- group kernels of the packed-byte shapes for 3-channel 8-bit stacks.
#include "k_batch.h"

namespace cp {
const void* u8_group_c3(int shape, bool absolute) { return u8_group_c<3>(shape, absolute); }
}  // namespace cp
