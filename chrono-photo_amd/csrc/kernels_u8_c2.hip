// kernels_u8_c2.hip -- 8-bit register-resident outlier kernels, 2 channel(s).
#include "kernels_u8.h"

namespace cp {

bool launch_u8_c2(const OutlierK& k, int shape, hipStream_t s) { return launch_u8_c<2>(k, shape, s); }

}  // namespace cp
