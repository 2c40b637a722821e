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
// Work split.  A workgroup owns one (frame, anchor, chunk of candidate rows, chunk of window
// rows): it stages the frame-f search region those rows reach (window chunk + candidate
// chunk - 1 rows, u16) and the frame-0 window rows (f32) in LDS.  Each candidate's sum is
// expanded as
//   sum (w - v)^2 = sum w^2 + sum v^2 - 2 sum w*v,
// all three exact integers.  sum w^2 is one constant for every frame and candidate: the
// frame-1 workgroups add it to a spare slot of the table, and the arg-min kernel adds it back.
// sum v^2 comes from an LDS table of squared-sample box sums per region row and candidate
// column (one sliding pass per row).  sum w*v is the inner loop: a lane owns a strip of K
// horizontally adjacent candidates (ox0..ox0+K-1, oy) over a sub-chunk of the rows, and the
// strip's region pixels slide through registers along dx, so one f64 FMA per term and each
// LDS sample serves K terms.  Products and sums of integers below 2^16 stay exact in f64
// while every partial sum is below 2^53, which the host guarantees by bounding the rows per
// lane.  Each lane adds sum v^2 - 2 sum w*v (modulo 2^64) for its K candidates to the
// per-frame table with 64-bit atomics: integer addition commutes, so the result does not
// depend on order, and the table holds the exact sums once the constant is added.
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
    int32_t cand_rows;        // candidate rows oy per workgroup
    int32_t cand_chunks;      // ceil((2s+1) / cand_rows); blockIdx.y = anchor * cand_chunks + chunk
    int32_t nsub;             // row sub-chunks per workgroup (one per group of strip lanes)
    int32_t spr;              // strips per offset row: ceil((2s+1) / K)
    int32_t strips_pad;       // spr * cand_rows rounded up to whole waves
    int32_t pitch_px;         // LDS region row pitch in pixels: spr * K + 2r
    unsigned long long* sums; // [n][(2s+1)^2], then the sum of squared window samples
    int32_t* offsets;         // [n][2]
};

// LDS layout of one workgroup for `rows` window rows and `crow` candidate rows: region
// samples (u16, rows + crow - 1 region rows of pitch_px pixels), window samples (f32,
// rows x (2r+1) pixels), squared-sample box sums (f64, one per region row and candidate
// column).
struct ShakeLds {
    uint32_t win_off, rq_off, bytes;
};
__host__ __device__ inline ShakeLds shake_lds(int32_t rows, int32_t crow, int32_t r, int32_t s, int32_t pitch_px,
                                             int32_t channels) {
    const uint64_t srows = (uint64_t)rows + crow - 1;
    const uint64_t region = srows * pitch_px * channels * 2;
    const uint64_t win_off = (region + 7) / 8 * 8;
    const uint64_t rq_off = (win_off + (uint64_t)rows * (2 * r + 1) * channels * 4 + 7) / 8 * 8;
    const uint64_t bytes = rq_off + srows * (2 * s + 1) * 8;
    const uint64_t cap = 0xFFFFFFFFull;
    return ShakeLds{(uint32_t)(win_off < cap ? win_off : cap), (uint32_t)(rq_off < cap ? rq_off : cap),
                    (uint32_t)(bytes < cap ? bytes : cap)};
}

// One dx step of a strip: load region pixel dx+K-1 into its register slot, then add the
// K products w(dx) * v(dx + j) for every channel.  U = dx mod K names the slots statically.
template <int C, int K, int U>
__device__ __forceinline__ void shake_step(const uint16_t* rr, const float* wr, int32_t dx, double (&v)[K][C],
                                           double (&acc)[K]) {
#pragma unroll
    for (int c = 0; c < C; ++c) v[(U + K - 1) % K][c] = (double)(uint32_t)rr[(dx + K - 1) * C + c];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double w = (double)wr[dx * C + c];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = __fma_rn(w, v[(U + j) % K][c], acc[j]);
    }
}

template <int C, int K, int U>
__device__ __forceinline__ void shake_steps(const uint16_t* rr, const float* wr, int32_t dx, double (&v)[K][C],
                                            double (&acc)[K]) {
    if constexpr (U < K) {
        shake_step<C, K, U>(rr, wr, dx + U, v, acc);
        shake_steps<C, K, U + 1>(rr, wr, dx, v, acc);
    }
}

// Remainder of a row (fewer than K steps, count uniform across the wave).
template <int C, int K, int U>
__device__ __forceinline__ void shake_tail(const uint16_t* rr, const float* wr, int32_t dx, int32_t left,
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
    extern __shared__ double lds_f64[];
    const uint32_t f = blockIdx.x + 1;
    const uint32_t a = blockIdx.y / (uint32_t)k.cand_chunks;
    const int32_t oc = (int32_t)(blockIdx.y - a * (uint32_t)k.cand_chunks);
    const int32_t r = k.r, s = k.s;
    const int32_t size = 2 * r + 1, ssize = 2 * s + 1;
    const int32_t c0 = (int32_t)blockIdx.z * k.chunk_rows;
    const int32_t crows = min(size, c0 + k.chunk_rows) - c0;
    const int32_t oy0 = oc * k.cand_rows;
    const int32_t orows = min(ssize, oy0 + k.cand_rows) - oy0;
    const int32_t srows = crows + orows - 1;  // region rows: the candidate rows over the chunk
    const int32_t pitch = k.pitch_px * C;     // samples per LDS region row
    const int32_t wl = size * C;              // samples per window row
    const ShakeLds lay = shake_lds(crows, orows, r, s, k.pitch_px, C);
    uint8_t* lds = reinterpret_cast<uint8_t*>(lds_f64);
    uint16_t* region = reinterpret_cast<uint16_t*>(lds);
    float* win = reinterpret_cast<float*>(lds + lay.win_off);
    double* rq = reinterpret_cast<double*>(lds + lay.rq_off);
    const int32_t cx = k.anchors[2 * a], cy = k.anchors[2 * a + 1];
    // Region pixel (x, y) is frame-f pixel (cx - s - r + x, cy - s - r + oy0 + c0 + y); the
    // pixels past 2r + 2s only pad the last strip of a row and are never summed.
    {
        // flat loops over (row, sample) with 8 loads in flight per lane
        constexpr int U = 8;
        const uint32_t nt = blockDim.x, tid = threadIdx.x;
        const uint64_t rstride = (uint64_t)k.stack_w * C;
        const uint32_t rlen = (uint32_t)((2 * r + 2 * s + 1) * C), rtotal = (uint32_t)(srows * pitch);
        const uint64_t fbase = k.pos_off[f] + ((uint64_t)(cy - s - r + oy0 + c0) * k.stack_w + (cx - s - r)) * C;
        for (uint32_t i0 = 0; i0 < rtotal; i0 += U * nt) {
            uint32_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * nt + tid, y = i / (uint32_t)pitch, e = i - y * (uint32_t)pitch;
                v[u] = i < rtotal && e < rlen ? load_sample<B>(k.stack, fbase + y * rstride + e) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * nt + tid;
                if (i < rtotal) region[i] = (uint16_t)v[u];
            }
        }
        const uint32_t wtotal = (uint32_t)(crows * wl);
        const uint64_t wbase = k.pos_off[0] + ((uint64_t)(cy - r + c0) * k.stack_w + (cx - r)) * C;
        double wsq = 0.0;
        for (uint32_t i0 = 0; i0 < wtotal; i0 += U * nt) {
            uint32_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * nt + tid, y = i / (uint32_t)wl, e = i - y * (uint32_t)wl;
                v[u] = i < wtotal ? load_sample<B>(k.stack, wbase + y * rstride + e) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * nt + tid;
                if (i < wtotal) {
                    const float w = (float)v[u];
                    win[i] = w;
                    wsq = __fma_rn((double)w, (double)w, wsq);
                }
            }
        }
        // sum of squared window samples, once per (anchor, window chunk): frame-independent
        if (f == 1 && oc == 0 && wsq != 0.0)
            atomicAdd(k.sums + (uint64_t)k.n * ssize * ssize, (unsigned long long)wsq);
    }
    __syncthreads();
    // rq[y][x] = sum over the window width of squared region samples starting at pixel x:
    // one lane per (region row, strip of K columns), sliding along the row
    for (int32_t t = threadIdx.x; t < srows * k.spr; t += blockDim.x) {
        const int32_t y = t / k.spr, x0 = (t - y * k.spr) * K;
        const uint16_t* row = region + y * pitch;
        double sum = 0.0;
        for (int32_t i = x0 * C; i < (x0 + size) * C; ++i) {
            const double v = (double)(uint32_t)row[i];
            sum = __fma_rn(v, v, sum);
        }
        rq[y * ssize + x0] = sum;
#pragma unroll
        for (int j = 1; j < K; ++j) {
            if (x0 + j >= ssize) break;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const double vin = (double)(uint32_t)row[(x0 + j - 1 + size) * C + c];
                const double vout = (double)(uint32_t)row[(x0 + j - 1) * C + c];
                sum = __fma_rn(vin, vin, sum);
                sum = __fma_rn(-vout, vout, sum);
            }
            rq[y * ssize + x0 + j] = sum;
        }
    }
    __syncthreads();
    const int32_t nstrips = k.spr * orows;
    const int32_t ntasks = k.strips_pad * k.nsub;
    const uint64_t noff = (uint64_t)ssize * ssize;
    for (int32_t t = threadIdx.x; t < ntasks; t += blockDim.x) {
        // strips_pad is a multiple of 64, so a wave shares its sub-chunk (and window reads)
        const int32_t sub = t / k.strips_pad, strip = t - sub * k.strips_pad;
        if (strip >= nstrips) continue;
        const int32_t oy = strip / k.spr, ox0 = (strip - oy * k.spr) * K;  // oy within the chunk
        const int32_t lo = sub * crows / k.nsub, hi = (sub + 1) * crows / k.nsub;
        double acc[K], sq[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = sq[j] = 0.0;
        for (int32_t y = lo; y < hi; ++y) {
            const uint16_t* rr = region + (oy + y) * pitch + ox0 * C;
            const float* wr = win + y * wl;
            double v[K][C];
#pragma unroll
            for (int u = 0; u < K - 1; ++u)
#pragma unroll
                for (int c = 0; c < C; ++c) v[u][c] = (double)(uint32_t)rr[u * C + c];
            int32_t dx = 0;
            for (; dx + K <= size; dx += K) shake_steps<C, K, 0>(rr, wr, dx, v, acc);
            shake_tail<C, K, 0>(rr, wr, dx, size - dx, v, acc);
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (ox0 + j < ssize) sq[j] += rq[(oy + y) * ssize + ox0 + j];
        }
        unsigned long long* out = k.sums + (uint64_t)f * noff + (uint64_t)(oy0 + oy) * ssize + ox0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (ox0 + j < ssize)
                atomicAdd(out + j, (unsigned long long)sq[j] - 2ull * (unsigned long long)acc[j]);
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
        unsigned long long raw = k.sums[(uint64_t)f * noff + o] + k.sums[(uint64_t)k.n * noff];
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
