// k_batch.h -- many outlier composites in a few persistent launches.
//
// Jobs (a stack, its frame selection and options) are grouped by tile shape and threshold
// kind; each group is one persistent launch.  Every job is cut into row tiles; a group's
// tiles form one work list, biggest jobs first.  A grid sized to the device's resident
// capacity loops: one lane of each workgroup takes the next item from a device counter, the
// workgroup stages that job's frame-offset table in LDS when the job changes, and runs the
// same tile body as the single-stack kernel (outlier_tile).  Results are bit-identical to
// one launch per stack: the per-pixel arithmetic is the same code.
#pragma once
#include "kernels_fast.h"

namespace cp {

struct BatchJobK {
    OutlierK k;        // pos_off points into the batch's device tables
    uint32_t tiles_x;  // tiles per output row
    int32_t shape;     // kShapes index, <= kBatchShapeMax
    uint32_t item0;    // first work item of this job
    uint32_t pad;
};

template <int B>
__global__ void batch_reset_kernel(unsigned int* counter) { *counter = 0u; }

// All jobs of one launch share the sample width B, the channel count C, the tile shape SH and
// the threshold kind ABS, so the kernel is the single-stack tile body with its own register
// budget.
template <int B, int C, int SH, bool ABS>
__global__ void __launch_bounds__(256) outlier_batch_kernel(const BatchJobK* __restrict__ jobs,
                                                            uint32_t njobs, uint32_t total,
                                                            unsigned int* __restrict__ counter) {
    constexpr int G = kShapes[SH].G, F = kShapes[SH].F;
    __shared__ TileShared<C, G, F> sh;
    __shared__ uint32_t s_item;
    uint32_t cur = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        // uniform across the workgroup; readfirstlane lets the job record live in SGPRs
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= total) break;
        uint32_t lo = 0, hi = njobs;  // last job whose first item is <= item
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].item0 <= item) lo = mid;
            else hi = mid;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const BatchJobK& J = jobs[lo];
        if (lo != cur) {
            load_offsets<B>(sh.off, G * F, J.k.pos_off, J.k.n);
            cur = lo;
        }
        __syncthreads();
        const uint32_t tile = item - J.item0;
        outlier_tile<B, C, G, F, ABS>(J.k, tile % J.tiles_x, tile / J.tiles_x, sh);
        __syncthreads();  // the tile's LDS and s_item are rewritten next iteration
    }
}

template <int B, int C, int SH, bool ABS>
inline int launch_batch_one(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter,
                            int cus, hipStream_t s) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, outlier_batch_kernel<B, C, SH, ABS>, 256, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const uint32_t grid = (uint32_t)(cus > 0 ? cus : 256) * (uint32_t)per_cu;
    hipLaunchKernelGGL(batch_reset_kernel<B>, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL((outlier_batch_kernel<B, C, SH, ABS>), dim3(grid), dim3(256), 0, s, jobs, njobs, total,
                       counter);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int B, int C>
inline int launch_batch_bc(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter,
                           int shape, bool absolute, int cus, hipStream_t s) {
    switch (shape * 2 + (absolute ? 1 : 0)) {
    case 0: return launch_batch_one<B, C, 0, false>(jobs, njobs, total, counter, cus, s);
    case 1: return launch_batch_one<B, C, 0, true>(jobs, njobs, total, counter, cus, s);
    case 2: return launch_batch_one<B, C, 1, false>(jobs, njobs, total, counter, cus, s);
    case 3: return launch_batch_one<B, C, 1, true>(jobs, njobs, total, counter, cus, s);
    case 4: return launch_batch_one<B, C, 2, false>(jobs, njobs, total, counter, cus, s);
    case 5: return launch_batch_one<B, C, 2, true>(jobs, njobs, total, counter, cus, s);
    case 6: return launch_batch_one<B, C, 3, false>(jobs, njobs, total, counter, cus, s);
    case 7: return launch_batch_one<B, C, 3, true>(jobs, njobs, total, counter, cus, s);
    case 8: return launch_batch_one<B, C, 4, false>(jobs, njobs, total, counter, cus, s);
    case 9: return launch_batch_one<B, C, 4, true>(jobs, njobs, total, counter, cus, s);
    case 10: return launch_batch_one<B, C, 5, false>(jobs, njobs, total, counter, cus, s);
    case 11: return launch_batch_one<B, C, 5, true>(jobs, njobs, total, counter, cus, s);
    case 12: return launch_batch_one<B, C, 6, false>(jobs, njobs, total, counter, cus, s);
    case 13: return launch_batch_one<B, C, 6, true>(jobs, njobs, total, counter, cus, s);
    case 14: return launch_batch_one<B, C, 7, false>(jobs, njobs, total, counter, cus, s);
    case 15: return launch_batch_one<B, C, 7, true>(jobs, njobs, total, counter, cus, s);
    default: return -1;
    }
}

// Defined per channel count for 8-bit stacks in kernels_batch_b1_c<C>.hip.
int launch_batch_b1_c3(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter, int shape,
                       bool absolute, int cus, hipStream_t s);
int launch_batch_b1_c4(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter, int shape,
                       bool absolute, int cus, hipStream_t s);

}  // namespace cp
