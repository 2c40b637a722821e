// kernels_fast_b1_c4.hip -- register-resident outlier kernel, 8-bit samples, 4 channels.
#include "kernels_fast.h"

namespace cp {

bool launch_fast_b1_c4(const OutlierK& k, int shape, hipStream_t s) { return launch_fast_c<1, 4>(k, shape, s); }

}  // namespace cp
