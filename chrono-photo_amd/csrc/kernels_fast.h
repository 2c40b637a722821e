// kernels_fast.h -- tile shapes of the register-resident outlier kernel and their launch.
//
// A shape (G, F) puts G lanes on each pixel and F frames in each lane (G * F >= N); a
// 256-lane workgroup then covers 256 / G pixels.  Frame counts fall into buckets; each bucket
// has up to two candidate shapes and the autotuner (cp_autotune) measures both on the device
// and records the faster one per (sample width, channels, bucket) in the context.
#pragma once
#include "k_outlier_fast.h"
#include "runtime.h"

namespace cp {

struct TileShape {
    int G, F;
};

constexpr int kNumShapes = 10;
constexpr TileShape kShapes[kNumShapes] = {{1, 16}, {1, 32}, {2, 16}, {2, 32}, {4, 16},
                                           {4, 32}, {8, 16}, {8, 32}, {4, 64}, {8, 64}};
constexpr int kNumBuckets = 6;
constexpr uint32_t kBucketMax[kNumBuckets] = {16, 32, 64, 128, 256, 512};
constexpr int kBucketCand[kNumBuckets][2] = {{0, -1}, {1, 2}, {3, 4}, {5, 6}, {7, 8}, {9, -1}};
// Shapes the batched kernels carry (F <= 32 keeps their register budget).
constexpr int kBatchShapeMax = 7;

inline int bucket_for(uint32_t n) {
    for (int b = 0; b < kNumBuckets; ++b)
        if (n <= kBucketMax[b]) return b;
    return -1;
}

template <int B, int C, int G, int F>
inline void launch_fast_shape(const OutlierK& k, hipStream_t s) {
    constexpr int P = 256 / G;
    const uint32_t tiles_x = (k.out_w + P - 1) / P;
    dim3 grid(tiles_x * k.out_h), block(256);
    if (k.th.absolute) hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, true>), grid, block, 0, s, k, tiles_x);
    else hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, false>), grid, block, 0, s, k, tiles_x);
}

template <int B, int C>
inline bool launch_fast_c(const OutlierK& k, int shape, hipStream_t s) {
    switch (shape) {
    case 0: launch_fast_shape<B, C, 1, 16>(k, s); return true;
    case 1: launch_fast_shape<B, C, 1, 32>(k, s); return true;
    case 2: launch_fast_shape<B, C, 2, 16>(k, s); return true;
    case 3: launch_fast_shape<B, C, 2, 32>(k, s); return true;
    case 4: launch_fast_shape<B, C, 4, 16>(k, s); return true;
    case 5: launch_fast_shape<B, C, 4, 32>(k, s); return true;
    case 6: launch_fast_shape<B, C, 8, 16>(k, s); return true;
    case 7: launch_fast_shape<B, C, 8, 32>(k, s); return true;
    case 8: launch_fast_shape<B, C, 4, 64>(k, s); return true;
    case 9: launch_fast_shape<B, C, 8, 64>(k, s); return true;
    default: return false;
    }
}

// One translation unit per (sample bytes, channels): kernels_fast_b<B>_c<C>.hip.
bool launch_fast_b1_c3(const OutlierK& k, int shape, hipStream_t s);
bool launch_fast_b1_c4(const OutlierK& k, int shape, hipStream_t s);
bool launch_fast_b2_c3(const OutlierK& k, int shape, hipStream_t s);
bool launch_fast_b2_c4(const OutlierK& k, int shape, hipStream_t s);

inline bool launch_outlier_fast(uint32_t bps, uint32_t channels, const OutlierK& k, int shape, hipStream_t s) {
    if (bps == 1 && channels == 3) return launch_fast_b1_c3(k, shape, s);
    if (bps == 1 && channels == 4) return launch_fast_b1_c4(k, shape, s);
    if (bps == 2 && channels == 3) return launch_fast_b2_c3(k, shape, s);
    if (bps == 2 && channels == 4) return launch_fast_b2_c4(k, shape, s);
    return false;
}

}  // namespace cp
