// kernels_fast_b2_c3.hip -- register-resident outlier kernel, 16-bit samples, 3 channels.
#include "kernels_fast.h"

namespace cp {

bool launch_fast_b2_c3(const OutlierK& k, int shape, hipStream_t s) { return launch_fast_c<2, 3>(k, shape, s); }

}  // namespace cp
