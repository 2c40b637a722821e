// kernels_batch_b1_c4.hip -- batched outlier kernels, 8-bit samples, 4 channels.
#include "k_batch.h"

namespace cp {

int launch_batch_b1_c4(const BatchJobK* jobs, uint32_t njobs, uint32_t total, int shape, bool absolute,
                       hipStream_t s) {
    return launch_batch_bc<1, 4>(jobs, njobs, total, shape, absolute, s);
}

}  // namespace cp
