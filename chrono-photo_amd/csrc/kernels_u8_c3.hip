// kernels_u8_c3.hip -- 8-bit register-resident outlier kernels, 3 channel(s).
#include "kernels_u8.h"

namespace cp {

bool launch_u8_c3(const OutlierK& k, int shape, hipStream_t s) { return launch_u8_c<3>(k, shape, s); }

}  // namespace cp
