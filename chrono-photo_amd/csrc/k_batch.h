// k_batch.h -- many outlier composites in a few launches, one per tile shape.
//
// The jobs that share a kernel shape (family, sample width, channels, shape, threshold kind)
// form one list of tiles, biggest jobs first, and one launch covers the list: workgroup b
// takes tile xcd_tile(b) (so each XCD runs a contiguous run of neighbouring tiles, as in the
// plain kernels), finds the tile's job by a binary search over the jobs' first tiles, stages
// that job's frame-offset table in LDS and runs the tile body of the plain kernel of the same
// shape.  Each shape is its own kernel with the register budget of its plain kernel; results
// are bit-identical to one launch per stack, the per-pixel arithmetic being the same code.
#pragma once
#include "kernels_fast.h"
#include "kernels_u8.h"

namespace cp {

struct BatchJobK {
    OutlierK k;        // device pointers into the batch's tables
    uint32_t tiles_x;  // tiles per output row
    uint32_t tile0;    // first tile of this job in its group's tile list
    uint32_t tiles;    // tiles_x * out_h
    uint32_t pad;
};

// p as a global-memory pointer: the address-space cast lets the compiler emit global (not
// flat) loads and stores through it.  Kernel arguments get that inference by themselves; the
// pointers of a job read from the job table do not, and flat accesses need 64-bit addresses
// in VGPRs and also count against the LDS wait counter.
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

// A job's arguments with every pointer known to be global memory.
__device__ __forceinline__ OutlierK global_args(const OutlierK& j) {
    OutlierK k = j;
    k.stack = as_global(k.stack);
    k.pos_off = as_global(k.pos_off);
    k.sample_mask = as_global(k.sample_mask);
    k.samp_off = as_global(k.samp_off);
    k.fade.values = as_global(k.fade.values);
    k.out_px = as_global(k.out_px);
    k.out_mask = as_global(k.out_mask);
    k.warnings = as_global(k.warnings);
    k.stats = as_global(k.stats);
    return k;
}

// The job of a tile: the last one whose first tile is <= tile.
__device__ __forceinline__ uint32_t group_job(const BatchJobK* __restrict__ jobs, uint32_t njobs, uint32_t tile) {
    uint32_t a = 0, b = njobs;
    while (b - a > 1) {
        const uint32_t m = (a + b) >> 1;
        if (jobs[m].tile0 <= tile) a = m;
        else b = m;
    }
    return a;
}

// The packed-byte kernel over a group of jobs (k_outlier_u8.h).
template <int C, int G, int F, int X, int FS, bool ABS, bool SORT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint_u8(C, F, X, FS, SORT))))
outlier_u8_group_kernel(const BatchJobK* __restrict__ jobs, uint32_t njobs) {
    __shared__ TileShared<C, G, F, X> sh;
    __shared__ uint64_t soff[FS > 0 ? G * (FS + 2) : 1];
    __shared__ float odist[keeps_odist<G, F, X, FS, SORT>() ? (kOdistSlots + 1) * TileShared<C, G, F, X>::P : 1];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const BatchJobK& J = jobs[group_job(jobs, njobs, tile)];
    const OutlierK k = global_args(J.k);
    load_offsets<1, F>(sh.off, G * F, k.pos_off, k.n);
    if constexpr (FS > 0) load_offsets<1, FS>(soff, G * FS, k.samp_off, k.n_sample);
    __syncthreads();
    const uint32_t t = tile - J.tile0;
    outlier_u8_tile<C, G, F, X, FS, ABS, SORT>(k, t % J.tiles_x, t / J.tiles_x, sh, soff, threadIdx.x, odist);
}

// The 16-bit-halves kernel over a group of jobs (k_outlier_fast.h).
template <int B, int C, int G, int F, int X, bool ABS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint(F, C, ABS, X))))
outlier_fast_group_kernel(const BatchJobK* __restrict__ jobs, uint32_t njobs) {
    __shared__ TileShared<C, G, F, X> sh;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const BatchJobK& J = jobs[group_job(jobs, njobs, tile)];
    const OutlierK k = global_args(J.k);
    const uint32_t t = tile - J.tile0, y = t / J.tiles_x;
    if constexpr (X == 2) {
        const uint32_t r = (uint32_t)((uintptr_t)k.stack + (uint64_t)y * k.stack_w * C * B) & 3u;
        load_offsets_pair<B, F>(sh.off, sh.aoff, sh.sft, G * F, k.pos_off, k.n, r);
    } else {
        load_offsets<B, F>(sh.off, G * F, k.pos_off, k.n);
    }
    __syncthreads();
    outlier_tile<B, C, G, F, X, ABS>(k, t % J.tiles_x, y, sh, threadIdx.x);
}

template <int C, int G, int F, int X, int FS, bool SORT = false>
constexpr const void* u8_group(bool absolute) {
    return absolute ? (const void*)&outlier_u8_group_kernel<C, G, F, X, FS, true, SORT>
                    : (const void*)&outlier_u8_group_kernel<C, G, F, X, FS, false, SORT>;
}

// The group kernel of packed-byte shape `shape` (kU8Shapes index), or nullptr.
template <int C>
const void* u8_group_c(int shape, bool absolute) {
    switch (shape) {
    case 0: return u8_group<C, 1, 16, 1, 0>(absolute);
    case 1: return u8_group<C, 1, 32, 1, 0>(absolute);
    case 2: return u8_group<C, 2, 32, 1, 0>(absolute);
    case 3: return u8_group<C, 4, 32, 1, 0>(absolute);
    case 4: return u8_group<C, 8, 32, 1, 0>(absolute);
    case 5: return u8_group<C, 16, 32, 1, 0>(absolute);
    case 6: return u8_group<C, 16, 32, 4, 0>(absolute);
    case 7: return u8_group<C, 32, 24, 4, 0>(absolute);
    case 8: return u8_group<C, 32, 32, 4, 0>(absolute);
    case 9: return u8_group<C, 32, 40, 4, 0>(absolute);
    case 10: return u8_group<C, 4, 32, 1, 8>(absolute);
    case 11: return u8_group<C, 8, 32, 1, 8>(absolute);
    case 12: return u8_group<C, 16, 32, 4, 8>(absolute);
    case 13: return u8_group<C, 32, 40, 4, 4>(absolute);
    case 14: return u8_group<C, 1, 16, 1, 0, true>(absolute);
    case 15: return u8_group<C, 1, 32, 1, 0, true>(absolute);
    case 16: return u8_group<C, 1, 64, 1, 0, true>(absolute);
    case 17: return u8_group<C, 1, 64, 1, 0>(absolute);
    default: return nullptr;
    }
}

// ONLY: the sample width the shape is instantiated for (0: both).
template <int B, int C, int G, int F, int X, int ONLY = 0>
constexpr const void* fast_group(bool absolute) {
    if constexpr (ONLY != 0 && ONLY != B) return nullptr;
    else
        return absolute ? (const void*)&outlier_fast_group_kernel<B, C, G, F, X, true>
                        : (const void*)&outlier_fast_group_kernel<B, C, G, F, X, false>;
}

// The group kernel of 16-bit-halves shape `shape` (kShapes index), or nullptr: 8-bit
// stacks use the one-pixel shapes, 16-bit stacks the shapes their frame-count buckets list
// and the one-pixel shapes those fall back to.
template <int B, int C>
const void* fast_group_bc(int shape, bool absolute) {
    switch (shape) {
    case 0: return fast_group<B, C, 1, 16, 1>(absolute);
    case 1: return fast_group<B, C, 1, 32, 1>(absolute);
    case 2: return fast_group<B, C, 2, 16, 1, 1>(absolute);
    case 3: return fast_group<B, C, 2, 32, 1>(absolute);
    case 4: return fast_group<B, C, 4, 16, 1, 1>(absolute);
    case 5: return fast_group<B, C, 4, 32, 1>(absolute);
    case 6: return fast_group<B, C, 8, 16, 1, 1>(absolute);
    case 7: return fast_group<B, C, 8, 32, 1>(absolute);
    case 8: return fast_group<B, C, 4, 64, 1, 1>(absolute);
    case 9: return fast_group<B, C, 8, 64, 1>(absolute);
    case 10: return fast_group<B, C, 1, 16, 2, 2>(absolute);
    case 11: return fast_group<B, C, 1, 32, 2, 2>(absolute);
    case 12: return fast_group<B, C, 2, 32, 2, 2>(absolute);
    case 13: return fast_group<B, C, 4, 32, 2, 2>(absolute);
    case 14: return fast_group<B, C, 8, 32, 2, 2>(absolute);
    default: return nullptr;
    }
}

// One translation unit per family and sample format: kernels_batch_*.hip.
const void* u8_group_c1(int shape, bool absolute);
const void* u8_group_c2(int shape, bool absolute);
const void* u8_group_c3(int shape, bool absolute);
const void* u8_group_c4(int shape, bool absolute);
const void* fast_group_b1_c3(int shape, bool absolute);
const void* fast_group_b1_c4(int shape, bool absolute);
const void* fast_group_b2_c3(int shape, bool absolute);
const void* fast_group_b2_c4(int shape, bool absolute);

// The group kernel for a plan (family, shape) over stacks of `bps` bytes per sample and
// `channels` channels, or nullptr when the plan has none.
inline const void* group_kernel(int family, int shape, uint32_t bps, uint32_t channels, bool absolute) {
    if (family == 2) {
        if (bps != 1) return nullptr;
        switch (channels) {
        case 1: return u8_group_c1(shape, absolute);
        case 2: return u8_group_c2(shape, absolute);
        case 3: return u8_group_c3(shape, absolute);
        case 4: return u8_group_c4(shape, absolute);
        default: return nullptr;
        }
    }
    if (family == 1) {
        if (bps == 1 && channels == 3) return fast_group_b1_c3(shape, absolute);
        if (bps == 1 && channels == 4) return fast_group_b1_c4(shape, absolute);
        if (bps == 2 && channels == 3) return fast_group_b2_c3(shape, absolute);
        if (bps == 2 && channels == 4) return fast_group_b2_c4(shape, absolute);
    }
    return nullptr;
}

// Pixels per tile of a plan's shape.
inline uint32_t group_tile_pixels(int family, int shape) {
    if (family == 2) return (uint32_t)(kU8Shapes[shape].X * 256 / kU8Shapes[shape].G);
    return (uint32_t)(kShapes[shape].X * 256 / kShapes[shape].G);
}

}  // namespace cp
