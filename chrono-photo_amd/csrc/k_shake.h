// k_shake.h -- camera-shake offsets (ShakeAnalyzer, src/shake.rs:190-385).
//
// For every frame f >= 1 and every candidate offset (ox, oy) in [-s, s]^2 the kernel sums
// (win - img)^2 over all anchors, window pixels (2r+1)^2 and channels, where win is the
// anchor window of frame 0 (fill_windows, :315-343) and img the frame-f pixel displaced by
// (ox, oy) (calc_diffs, :345-385).  The offset with the smallest sum wins; ties go to the
// first index in row-major (oy, ox) order (min_by_key, :294-297).
//
// u8 sums wrap in i32 like the reference (release build, overflow-checks = false); they
// are accumulated in u32, which is the same value modulo 2^32.  u16 stacks accumulate in
// int64.
//
// One workgroup per (frame, anchor) pair stages the anchor's search region of frame f in
// LDS; each lane owns candidate offsets and streams the frame-0 window from cache
// (wave-uniform addresses).  Per-anchor partial sums go to a per-frame table with 64-bit
// atomics (integer addition commutes, so the result does not depend on order); a second
// kernel takes the arg-min.
#pragma once
#include "device_common.h"

namespace cp {

struct ShakeK {
    const uint8_t* stack;
    const uint64_t* pos_off;  // element offset of each frame position (no crop)
    uint32_t stack_w;
    uint32_t n;
    const int32_t* anchors;   // device, 2 per anchor
    uint32_t n_anchors;
    int32_t r, s;
    unsigned long long* sums; // [n][(2s+1)^2]
    int32_t* offsets;         // [n][2]
};

template <int B, int C>
__global__ void __launch_bounds__(256) shake_ssd_kernel(ShakeK k) {
    extern __shared__ uint16_t region[];  // (2r+2s+1)^2 * C samples of frame f
    const uint32_t f = blockIdx.x + 1;
    const uint32_t a = blockIdx.y;
    const int32_t r = k.r, s = k.s;
    const int32_t size = 2 * r + 1, ssize = 2 * s + 1, rsize = size + 2 * s;
    const int32_t cx = k.anchors[2 * a], cy = k.anchors[2 * a + 1];
    // region origin: (cx - s - r, cy - s - r)
    const int32_t rx0 = cx - s - r, ry0 = cy - s - r;
    const uint64_t fbase = k.pos_off[f];
    const uint64_t wbase = k.pos_off[0];
    const int32_t rlen = rsize * rsize * C;
    for (int32_t i = threadIdx.x; i < rlen; i += blockDim.x) {
        int32_t ch = i % C, p = i / C;
        int32_t xx = rx0 + p % rsize, yy = ry0 + p / rsize;
        region[i] = (uint16_t)load_sample<B>(k.stack, fbase + ((uint64_t)yy * k.stack_w + xx) * C + ch);
    }
    __syncthreads();
    const int32_t noff = ssize * ssize;
    for (int32_t o = threadIdx.x; o < noff; o += blockDim.x) {
        const int32_t ox = o % ssize, oy = o / ssize;
        uint64_t acc64 = 0;
        uint32_t acc32 = 0;
        for (int32_t dy = 0; dy < size; ++dy) {
            const uint64_t wrow = wbase + ((uint64_t)(cy + dy - r) * k.stack_w + (cx - r)) * C;
            const uint16_t* srow = region + ((oy + dy) * rsize + ox) * C;
            for (int32_t dx = 0; dx < size; ++dx) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) {
                    int32_t w = (int32_t)load_sample<B>(k.stack, wrow + dx * C + ch);
                    int32_t v = (int32_t)srow[dx * C + ch];
                    int32_t d = w - v;
                    if constexpr (B == 1) acc32 += (uint32_t)(d * d);
                    else acc64 += (uint64_t)((int64_t)d * d);
                }
            }
        }
        unsigned long long add = B == 1 ? (unsigned long long)acc32 : (unsigned long long)acc64;
        atomicAdd(&k.sums[(uint64_t)f * noff + o], add);
    }
}

template <int B>
__global__ void __launch_bounds__(256) shake_argmin_kernel(ShakeK k) {
    const uint32_t f = blockIdx.x;
    const int32_t ssize = 2 * k.s + 1, noff = ssize * ssize;
    __shared__ long long best_v[256];
    __shared__ int32_t best_i[256];
    long long bv = 0;
    int32_t bi = -1;
    for (int32_t o = threadIdx.x; o < noff; o += blockDim.x) {
        unsigned long long raw = k.sums[(uint64_t)f * noff + o];
        long long v = B == 1 ? (long long)(int32_t)(uint32_t)raw : (long long)raw;
        if (bi < 0 || v < bv) { bv = v; bi = o; }
    }
    best_v[threadIdx.x] = bv;
    best_i[threadIdx.x] = bi;
    __syncthreads();
    for (int32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int32_t)threadIdx.x < w) {
            int32_t oi = best_i[threadIdx.x + w];
            long long ov = best_v[threadIdx.x + w];
            int32_t mi = best_i[threadIdx.x];
            long long mv = best_v[threadIdx.x];
            if (oi >= 0 && (mi < 0 || ov < mv || (ov == mv && oi < mi))) {
                best_i[threadIdx.x] = oi;
                best_v[threadIdx.x] = ov;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int32_t b = f == 0 ? -1 : best_i[0];
        if (f == 0 || b < 0) {
            k.offsets[2 * f] = 0;
            k.offsets[2 * f + 1] = 0;
        } else {
            k.offsets[2 * f] = b % ssize - k.s;
            k.offsets[2 * f + 1] = b / ssize - k.s;
        }
    }
}

}  // namespace cp
