// k_shake.h -- camera-shake offsets (ShakeAnalyzer, src/shake.rs:190-385).
//
// For every frame f >= 1 and every candidate offset (ox, oy) in [-s, s]^2 the kernel sums
// (win - img)^2 over all anchors, window pixels (2r+1)^2 and channels, where win is the
// anchor window of frame 0 (fill_windows, :315-343) and img the frame-f pixel displaced by
// (ox, oy) (calc_diffs, :345-385).  The offset with the smallest sum wins; ties go to the
// first index in row-major (oy, ox) order (min_by_key, :294-297).
//
// u8 sums wrap in i32 like the reference (release build, overflow-checks = false): the
// kernels sum exactly and the arg-min reads each total modulo 2^32, which is that wrapped
// value.  u16 sums are exact 64-bit integers.
//
// Work split.  A workgroup owns one (frame, anchor, chunk of window rows): it stages the
// frame-f search region those rows reach (chunk + 2s rows) and the frame-0 window rows in
// LDS.  A lane owns a strip of K horizontally adjacent candidate offsets (ox0..ox0+K-1, oy)
// over a sub-chunk of the rows; the strip's K region pixels slide through registers along
// dx, so each region sample read from LDS serves K terms and each window sample (the same
// address for the whole wave) serves K terms as well.  Terms are summed in f64: (w - v)^2
// of two integers below 2^16 is exact, and so is every partial sum below 2^53, which the
// host guarantees by bounding the rows per lane.  Each lane adds its K exact sums to the
// per-frame table with 64-bit atomics (integer addition commutes, so the result does not
// depend on order); a second kernel takes the arg-min.
#pragma once
#include "device_common.h"

namespace cp {

constexpr int kShakeStrip = 5;  // candidate offsets per lane (K)

struct ShakeK {
    const uint8_t* stack;
    const uint64_t* pos_off;  // element offset of each frame position (no crop)
    uint32_t stack_w;
    uint32_t n;
    const int32_t* anchors;   // device, 2 per anchor
    uint32_t n_anchors;
    int32_t r, s;
    int32_t chunk_rows;       // window rows per workgroup (blockIdx.z chunks them)
    int32_t nsub;             // row sub-chunks per workgroup (one per group of strip lanes)
    int32_t spr;              // strips per offset row: ceil((2s+1) / K)
    int32_t strips_pad;       // spr * (2s+1) rounded up to whole waves
    int32_t pitch_px;         // LDS region row pitch in pixels: spr * K + 2r
    unsigned long long* sums; // [n][(2s+1)^2]
    int32_t* offsets;         // [n][2]
};

// LDS bytes of one workgroup for a chunk of `rows` window rows.
inline size_t shake_lds_bytes(int32_t rows, int32_t r, int32_t s, int32_t pitch_px, int32_t channels) {
    const size_t srows = (size_t)rows + 2 * s;
    return (srows * pitch_px + (size_t)rows * (2 * r + 1)) * channels * sizeof(uint16_t);
}

// One dx step of a strip: load region pixel dx+K-1 into its register slot, then add the
// K terms (w(dx) - v(dx + j))^2 for every channel.  U = dx mod K names the slots statically.
template <int C, int K, int U>
__device__ __forceinline__ void shake_step(const uint16_t* rr, const uint16_t* wr, int32_t dx, double (&v)[K][C],
                                           double (&acc)[K]) {
#pragma unroll
    for (int c = 0; c < C; ++c) v[(U + K - 1) % K][c] = (double)(uint32_t)rr[(dx + K - 1) * C + c];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double w = (double)(uint32_t)wr[dx * C + c];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double d = w - v[(U + j) % K][c];
            acc[j] = __fma_rn(d, d, acc[j]);
        }
    }
}

template <int C, int K, int U>
__device__ __forceinline__ void shake_steps(const uint16_t* rr, const uint16_t* wr, int32_t dx, double (&v)[K][C],
                                            double (&acc)[K]) {
    if constexpr (U < K) {
        shake_step<C, K, U>(rr, wr, dx + U, v, acc);
        shake_steps<C, K, U + 1>(rr, wr, dx, v, acc);
    }
}

// Remainder of a row (fewer than K steps, count uniform across the wave).
template <int C, int K, int U>
__device__ __forceinline__ void shake_tail(const uint16_t* rr, const uint16_t* wr, int32_t dx, int32_t left,
                                           double (&v)[K][C], double (&acc)[K]) {
    if constexpr (U < K - 1) {
        if (U < left) {
            shake_step<C, K, U>(rr, wr, dx + U, v, acc);
            shake_tail<C, K, U + 1>(rr, wr, dx, left, v, acc);
        }
    }
}

template <int B, int C>
__global__ void __launch_bounds__(256) shake_ssd_kernel(ShakeK k) {
    constexpr int K = kShakeStrip;
    extern __shared__ uint16_t lds[];
    const uint32_t f = blockIdx.x + 1;
    const uint32_t a = blockIdx.y;
    const int32_t r = k.r, s = k.s;
    const int32_t size = 2 * r + 1, ssize = 2 * s + 1;
    const int32_t c0 = (int32_t)blockIdx.z * k.chunk_rows;
    const int32_t crows = min(size, c0 + k.chunk_rows) - c0;
    const int32_t srows = crows + ssize - 1;  // region rows: offset rows 0..2s over the chunk
    const int32_t pitch = k.pitch_px * C;     // samples per LDS region row
    const int32_t wl = size * C;              // samples per window row
    uint16_t* region = lds;
    uint16_t* win = lds + srows * pitch;
    const int32_t cx = k.anchors[2 * a], cy = k.anchors[2 * a + 1];
    // Region pixel (x, y) is frame-f pixel (cx - s - r + x, cy - s - r + c0 + y); the
    // pixels past 2r + 2s only pad the last strip of a row and are never summed.
    {
        const int32_t rlen = (2 * r + 2 * s + 1) * C;
        const uint64_t fbase = k.pos_off[f] + ((uint64_t)(cy - s - r + c0) * k.stack_w + (cx - s - r)) * C;
#pragma unroll 2
        for (int32_t y = 0; y < srows; ++y) {
            const uint64_t row = fbase + (uint64_t)y * k.stack_w * C;
            for (int32_t e = threadIdx.x; e < pitch; e += blockDim.x)
                region[y * pitch + e] = e < rlen ? (uint16_t)load_sample<B>(k.stack, row + e) : (uint16_t)0;
        }
        const uint64_t wbase = k.pos_off[0] + ((uint64_t)(cy - r + c0) * k.stack_w + (cx - r)) * C;
#pragma unroll 2
        for (int32_t y = 0; y < crows; ++y) {
            const uint64_t row = wbase + (uint64_t)y * k.stack_w * C;
            for (int32_t e = threadIdx.x; e < wl; e += blockDim.x)
                win[y * wl + e] = (uint16_t)load_sample<B>(k.stack, row + e);
        }
    }
    __syncthreads();
    const int32_t nstrips = k.spr * ssize;
    const int32_t ntasks = k.strips_pad * k.nsub;
    const uint64_t noff = (uint64_t)ssize * ssize;
    for (int32_t t = threadIdx.x; t < ntasks; t += blockDim.x) {
        // strips_pad is a multiple of 64, so a wave shares its sub-chunk (and window reads)
        const int32_t sub = t / k.strips_pad, strip = t - sub * k.strips_pad;
        if (strip >= nstrips) continue;
        const int32_t oy = strip / k.spr, ox0 = (strip - oy * k.spr) * K;
        const int32_t lo = sub * crows / k.nsub, hi = (sub + 1) * crows / k.nsub;
        double acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = 0.0;
        for (int32_t y = lo; y < hi; ++y) {
            const uint16_t* rr = region + (oy + y) * pitch + ox0 * C;
            const uint16_t* wr = win + y * wl;
            double v[K][C];
#pragma unroll
            for (int u = 0; u < K - 1; ++u)
#pragma unroll
                for (int c = 0; c < C; ++c) v[u][c] = (double)(uint32_t)rr[u * C + c];
            int32_t dx = 0;
            for (; dx + K <= size; dx += K) shake_steps<C, K, 0>(rr, wr, dx, v, acc);
            shake_tail<C, K, 0>(rr, wr, dx, size - dx, v, acc);
        }
        unsigned long long* out = k.sums + (uint64_t)f * noff + (uint64_t)oy * ssize + ox0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (ox0 + j < ssize) atomicAdd(out + j, (unsigned long long)acc[j]);
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
