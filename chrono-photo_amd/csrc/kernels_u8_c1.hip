// kernels_u8_c1.hip -- 8-bit register-resident outlier kernels, 1 channel(s).
#include "kernels_u8.h"

namespace cp {

bool launch_u8_c1(const OutlierK& k, int shape, hipStream_t s) { return launch_u8_c<1>(k, shape, s); }

}  // namespace cp
