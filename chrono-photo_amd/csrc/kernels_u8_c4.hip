// kernels_u8_c4.hip -- 8-bit register-resident outlier kernels, 4 channel(s).
#include "kernels_u8.h"

namespace cp {

bool launch_u8_c4(const OutlierK& k, int shape, hipStream_t s) { return launch_u8_c<4>(k, shape, s); }

}  // namespace cp
