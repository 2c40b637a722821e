// kernels_fast_b1_c3.hip -- register-resident outlier kernel, 8-bit samples, 3 channels.
#include "kernels_fast.h"

namespace cp {

bool launch_fast_b1_c3(const OutlierK& k, int shape, hipStream_t s) { return launch_fast_c<1, 3>(k, shape, s); }

}  // namespace cp
