// k_batch.h -- many outlier composites in one persistent launch.
//
// Every job (a stack, its frame selection and options) is cut into row tiles of its own
// tile shape; the jobs' tiles form one work list, biggest jobs first.  A grid sized to the
// device's resident capacity loops: one lane of each workgroup takes the next item from a
// device counter, the workgroup stages that job's frame-offset table in LDS when the job
// changes, and runs the same tile body as the single-stack kernel (outlier_tile), selected
// per job from the shapes with F <= 32 and both threshold kinds.  Results are bit-identical
// to one launch per stack: the per-pixel arithmetic is the same code.
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

template <int B, int C, int SH, bool ABS>
__device__ __forceinline__ void batch_tile(const OutlierK& k, uint32_t bx, uint32_t y, unsigned char* smem) {
    constexpr int G = kShapes[SH].G, F = kShapes[SH].F;
    outlier_tile<B, C, G, F, ABS>(k, bx, y, *reinterpret_cast<TileShared<C, G, F>*>(smem));
}

// Frame positions covered by a tile of shape sh (G * F).
__device__ __forceinline__ uint32_t shape_frames(int sh) {
    constexpr uint32_t gf[kBatchShapeMax + 1] = {
        kShapes[0].G * kShapes[0].F, kShapes[1].G * kShapes[1].F, kShapes[2].G * kShapes[2].F,
        kShapes[3].G * kShapes[3].F, kShapes[4].G * kShapes[4].F, kShapes[5].G * kShapes[5].F,
        kShapes[6].G * kShapes[6].F, kShapes[7].G * kShapes[7].F};
    return gf[sh];
}

// LDS of the largest tile shape the batch kernel carries.
template <int C, int SH>
constexpr size_t batch_smem() {
    if constexpr (SH < 0) return 0;
    else {
        constexpr size_t a = sizeof(TileShared<C, kShapes[SH].G, kShapes[SH].F>);
        constexpr size_t b = batch_smem<C, SH - 1>();
        return a > b ? a : b;
    }
}

template <int B>
__global__ void batch_reset_kernel(unsigned int* counter) { *counter = 0u; }

template <int B, int C>
__global__ void __launch_bounds__(256) outlier_batch_kernel(const BatchJobK* __restrict__ jobs,
                                                            uint32_t njobs, uint32_t total,
                                                            unsigned int* __restrict__ counter) {
    // every TileShared starts with its frame-offset table
    __shared__ __align__(16) unsigned char smem[batch_smem<C, kBatchShapeMax>()];
    uint64_t* const s_off = reinterpret_cast<uint64_t*>(smem);
    __shared__ uint32_t s_item;
    uint32_t cur = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= total) break;  // every lane of the workgroup sees the same item
        uint32_t lo = 0, hi = njobs;  // last job whose first item is <= item
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].item0 <= item) lo = mid;
            else hi = mid;
        }
        const BatchJobK& J = jobs[lo];
        if (lo != cur) {
            const uint32_t nf = shape_frames(J.shape);
            load_offsets<B>(s_off, nf, J.k.pos_off, J.k.n);
            cur = lo;
        }
        __syncthreads();
        const uint32_t tile = item - J.item0;
        const uint32_t bx = tile % J.tiles_x, y = tile / J.tiles_x;
        switch (J.shape * 2 + (J.k.th.absolute ? 1 : 0)) {
        case 0: batch_tile<B, C, 0, false>(J.k, bx, y, smem); break;
        case 1: batch_tile<B, C, 0, true>(J.k, bx, y, smem); break;
        case 2: batch_tile<B, C, 1, false>(J.k, bx, y, smem); break;
        case 3: batch_tile<B, C, 1, true>(J.k, bx, y, smem); break;
        case 4: batch_tile<B, C, 2, false>(J.k, bx, y, smem); break;
        case 5: batch_tile<B, C, 2, true>(J.k, bx, y, smem); break;
        case 6: batch_tile<B, C, 3, false>(J.k, bx, y, smem); break;
        case 7: batch_tile<B, C, 3, true>(J.k, bx, y, smem); break;
        case 8: batch_tile<B, C, 4, false>(J.k, bx, y, smem); break;
        case 9: batch_tile<B, C, 4, true>(J.k, bx, y, smem); break;
        case 10: batch_tile<B, C, 5, false>(J.k, bx, y, smem); break;
        case 11: batch_tile<B, C, 5, true>(J.k, bx, y, smem); break;
        case 12: batch_tile<B, C, 6, false>(J.k, bx, y, smem); break;
        case 13: batch_tile<B, C, 6, true>(J.k, bx, y, smem); break;
        case 14: batch_tile<B, C, 7, false>(J.k, bx, y, smem); break;
        default: batch_tile<B, C, 7, true>(J.k, bx, y, smem); break;
        }
        __syncthreads();  // the tile's LDS and s_item are rewritten next iteration
    }
}

// Defined per channel count for 8-bit stacks in kernels_batch_b1_c<C>.hip.
int launch_batch_b1_c3(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter,
                       int cus, hipStream_t s);
int launch_batch_b1_c4(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter,
                       int cus, hipStream_t s);

template <int B, int C>
inline int launch_batch_bc(const BatchJobK* jobs, uint32_t njobs, uint32_t total, unsigned int* counter,
                           int cus, hipStream_t s) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, outlier_batch_kernel<B, C>, 256, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const uint32_t grid = (uint32_t)(cus > 0 ? cus : 256) * (uint32_t)per_cu;
    hipLaunchKernelGGL(batch_reset_kernel<B>, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL((outlier_batch_kernel<B, C>), dim3(grid), dim3(256), 0, s, jobs, njobs, total, counter);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace cp
