// k_batch.h -- many outlier composites in a few launches.
//
// Jobs (a stack, its frame selection and options) are grouped by sample width, channel
// count, tile shape and threshold kind; each group is one launch whose grid covers the tiles
// of all its jobs, biggest jobs first.  Block b (renumbered XCD-contiguously, as in the
// single-stack kernel) finds its job by a binary search over the jobs' first tiles, stages
// that job's frame-offset table in LDS and runs the same tile body as the single-stack
// kernel (outlier_tile).  Results are bit-identical to one launch per stack: the per-pixel
// arithmetic is the same code.  A batch's launches are captured once into a HIP graph.
#pragma once
#include "kernels_fast.h"

namespace cp {

struct BatchJobK {
    OutlierK k;        // pos_off points into the batch's device tables
    uint32_t tiles_x;  // tiles per output row
    int32_t shape;     // kShapes index, <= kBatchShapeMax
    uint32_t tile0;    // first tile of this job in the group's grid
    uint32_t tiles;    // tiles_x * out_h
};

template <int B, int C, int SH, bool ABS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint(kShapes[SH].F, C, ABS))))
outlier_batch_kernel(const BatchJobK* __restrict__ jobs, uint32_t njobs) {
    constexpr int G = kShapes[SH].G, F = kShapes[SH].F;
    __shared__ TileShared<C, G, F> sh;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    uint32_t lo = 0, hi = njobs;  // last job whose first tile is <= tile
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid;
    }
    const BatchJobK& J = jobs[lo];
    const OutlierK k = J.k;
    load_offsets<B, F>(sh.off, G * F, k.pos_off, k.n);
    __syncthreads();
    const uint32_t t = tile - J.tile0;
    outlier_tile<B, C, G, F, 1, ABS>(k, t % J.tiles_x, t / J.tiles_x, sh);
}

template <int B, int C, int SH, bool ABS>
inline int launch_batch_one(const BatchJobK* jobs, uint32_t njobs, uint32_t total, hipStream_t s) {
    hipLaunchKernelGGL((outlier_batch_kernel<B, C, SH, ABS>), dim3(total), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int B, int C>
inline int launch_batch_bc(const BatchJobK* jobs, uint32_t njobs, uint32_t total, int shape, bool absolute,
                           hipStream_t s) {
    switch (shape * 2 + (absolute ? 1 : 0)) {
    case 0: return launch_batch_one<B, C, 0, false>(jobs, njobs, total, s);
    case 1: return launch_batch_one<B, C, 0, true>(jobs, njobs, total, s);
    case 2: return launch_batch_one<B, C, 1, false>(jobs, njobs, total, s);
    case 3: return launch_batch_one<B, C, 1, true>(jobs, njobs, total, s);
    case 4: return launch_batch_one<B, C, 2, false>(jobs, njobs, total, s);
    case 5: return launch_batch_one<B, C, 2, true>(jobs, njobs, total, s);
    case 6: return launch_batch_one<B, C, 3, false>(jobs, njobs, total, s);
    case 7: return launch_batch_one<B, C, 3, true>(jobs, njobs, total, s);
    case 8: return launch_batch_one<B, C, 4, false>(jobs, njobs, total, s);
    case 9: return launch_batch_one<B, C, 4, true>(jobs, njobs, total, s);
    case 10: return launch_batch_one<B, C, 5, false>(jobs, njobs, total, s);
    case 11: return launch_batch_one<B, C, 5, true>(jobs, njobs, total, s);
    case 12: return launch_batch_one<B, C, 6, false>(jobs, njobs, total, s);
    case 13: return launch_batch_one<B, C, 6, true>(jobs, njobs, total, s);
    case 14: return launch_batch_one<B, C, 7, false>(jobs, njobs, total, s);
    case 15: return launch_batch_one<B, C, 7, true>(jobs, njobs, total, s);
    default: return -1;
    }
}

// Defined per channel count for 8-bit stacks in kernels_batch_b1_c<C>.hip.
int launch_batch_b1_c3(const BatchJobK* jobs, uint32_t njobs, uint32_t total, int shape, bool absolute,
                       hipStream_t s);
int launch_batch_b1_c4(const BatchJobK* jobs, uint32_t njobs, uint32_t total, int shape, bool absolute,
                       hipStream_t s);

}  // namespace cp
