// kernels_fast_u16.hip -- instantiations of the register-resident outlier kernel for
// 16-bit stacks (RGB and RGBA; other channel counts use the generic kernel).
#include "kernels_fast.h"

namespace cp {

bool launch_outlier_fast_u16(const cp_stack* st, const OutlierK& k, hipStream_t s) {
    if (st->channels == 3) return launch_fast_c<2, 3>(k, s);
    if (st->channels == 4) return launch_fast_c<2, 4>(k, s);
    return false;
}

}  // namespace cp
