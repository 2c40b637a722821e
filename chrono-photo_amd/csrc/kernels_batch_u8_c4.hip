// kernels_batch_u8_c4.hip -- group kernels of the packed-byte shapes for 4-channel 8-bit stacks.
#include "k_batch.h"

namespace cp {
const void* u8_group_c4(int shape, bool absolute) { return u8_group_c<4>(shape, absolute); }
}  // namespace cp
