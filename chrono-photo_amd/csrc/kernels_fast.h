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
    int G, F, X;
};

// Shapes 10-14 put two pixels in each lane (16-bit stacks with 3 or 4 channels): one 12- or
// 16-byte load per frame instead of one per pixel, which is what lets a wave's loads reach
// HBM speed; narrow_shape gives the one-pixel shape with the same blocks.
constexpr int kNumShapes = 15;
constexpr TileShape kShapes[kNumShapes] = {{1, 16, 1}, {1, 32, 1}, {2, 16, 1}, {2, 32, 1}, {4, 16, 1},
                                           {4, 32, 1}, {8, 16, 1}, {8, 32, 1}, {4, 64, 1}, {8, 64, 1},
                                           {1, 16, 2}, {1, 32, 2}, {2, 32, 2}, {4, 32, 2}, {8, 32, 2}};
// Frame-count buckets of the tuned kernel choice.  A choice is a shape of this kernel
// (kShapes index) or, from kU8Base on, a shape of the 8-bit packed-byte kernel (kU8Shapes
// index + kU8Base, kernels_u8.h); each bucket has up to kMaxCand candidates, the first being
// the default, and cp_autotune keeps the fastest per (sample width, channels, bucket).
constexpr int kNumBuckets = 9;
constexpr uint32_t kBucketMax[kNumBuckets] = {16, 32, 64, 128, 256, 512, 768, 1024, 1280};
constexpr int kU8Base = 64;
constexpr int kMaxCand = 5;
// [sample bytes - 1][channels >= 3][bucket][candidate]; -1: none
constexpr int kBucketCand[2][2][kNumBuckets][kMaxCand] = {
    // 8-bit, 1-2 channels: the packed-byte kernel only
    {{{kU8Base + 0, kU8Base + 14, -1, -1, -1}, {kU8Base + 1, kU8Base + 15, -1, -1, -1},
      {kU8Base + 2, kU8Base + 16, kU8Base + 17, -1, -1}, {kU8Base + 3, -1, -1, -1, -1},
      {kU8Base + 4, -1, -1, -1, -1}, {kU8Base + 6, kU8Base + 5, -1, -1, -1}, {kU8Base + 7, -1, -1, -1, -1},
      {kU8Base + 8, -1, -1, -1, -1}, {kU8Base + 9, -1, -1, -1, -1}},
     // 8-bit, 3-4 channels
     {{0, kU8Base + 0, kU8Base + 14, -1, -1}, {1, 2, kU8Base + 1, kU8Base + 15, -1},
      {3, 4, kU8Base + 2, kU8Base + 16, kU8Base + 17}, {5, 6, kU8Base + 3, -1, -1},
      {7, 8, kU8Base + 4, -1, -1}, {9, kU8Base + 6, kU8Base + 5, -1, -1}, {kU8Base + 7, -1, -1, -1, -1},
      {kU8Base + 8, -1, -1, -1, -1}, {kU8Base + 9, -1, -1, -1, -1}}},
    // 16-bit: 1-2 channels have no register-resident kernel; 3-4 channels up to 512 frames
    {{{-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1},
      {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1},
      {-1, -1, -1, -1, -1}},
     {{10, 0, -1, -1, -1}, {11, 1, -1, -1, -1}, {12, 3, -1, -1, -1}, {13, 5, -1, -1, -1}, {14, 7, -1, -1, -1},
      {9, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}}}};
// Default choice for relative thresholds where it differs from the bucket's first candidate
// (-1: the first candidate): up to 64 frames of an 8-bit stack, one sort of a lane's bytes
// gives the median and both quartiles at once, where the byte search runs three searches
// (64 × 1080p RGB8, rel/2/4: 0.339 ms against 0.428 ms for the best byte-search tile).
constexpr int kRelDefault[2][2][kNumBuckets] = {
    {{kU8Base + 14, kU8Base + 15, kU8Base + 16, -1, -1, -1, -1, -1, -1},
     {kU8Base + 14, kU8Base + 15, kU8Base + 16, -1, -1, -1, -1, -1, -1}},
    {{-1, -1, -1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1, -1, -1}}};

inline int bucket_for(uint32_t n) {
    for (int b = 0; b < kNumBuckets; ++b)
        if (n <= kBucketMax[b]) return b;
    return -1;
}

inline int narrow_shape(int shape) {
    for (int i = 0; i < kNumShapes; ++i)
        if (kShapes[i].X == 1 && kShapes[i].G == kShapes[shape].G && kShapes[i].F == kShapes[shape].F) return i;
    return -1;
}

template <int B, int C, int G, int F, int X>
inline void launch_fast_shape(const OutlierK& k, hipStream_t s) {
    constexpr int P = X * 256 / G;
    const uint32_t tiles_x = (k.out_w + P - 1) / P;
    dim3 grid(tiles_x * k.out_h), block(256);
    if (k.th.absolute) hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, X, true>), grid, block, 0, s, k, tiles_x);
    else hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, X, false>), grid, block, 0, s, k, tiles_x);
}

template <int B, int C>
inline bool launch_fast_c(const OutlierK& k, int shape, hipStream_t s) {
    switch (shape) {
    case 0: launch_fast_shape<B, C, 1, 16, 1>(k, s); return true;
    case 1: launch_fast_shape<B, C, 1, 32, 1>(k, s); return true;
    case 2: launch_fast_shape<B, C, 2, 16, 1>(k, s); return true;
    case 3: launch_fast_shape<B, C, 2, 32, 1>(k, s); return true;
    case 4: launch_fast_shape<B, C, 4, 16, 1>(k, s); return true;
    case 5: launch_fast_shape<B, C, 4, 32, 1>(k, s); return true;
    case 6: launch_fast_shape<B, C, 8, 16, 1>(k, s); return true;
    case 7: launch_fast_shape<B, C, 8, 32, 1>(k, s); return true;
    case 8: launch_fast_shape<B, C, 4, 64, 1>(k, s); return true;
    case 9: launch_fast_shape<B, C, 8, 64, 1>(k, s); return true;
    default: break;
    }
    if constexpr (B == 2) {
        switch (shape) {
        case 10: launch_fast_shape<B, C, 1, 16, 2>(k, s); return true;
        case 11: launch_fast_shape<B, C, 1, 32, 2>(k, s); return true;
        case 12: launch_fast_shape<B, C, 2, 32, 2>(k, s); return true;
        case 13: launch_fast_shape<B, C, 4, 32, 2>(k, s); return true;
        case 14: launch_fast_shape<B, C, 8, 32, 2>(k, s); return true;
        default: break;
        }
    }
    return false;
}

// Whether a two-pixel tile may run over this composite: the stack starts on an 8-byte
// boundary (4-channel pairs are loaded as they lie, 3-channel pairs realigned from dword
// boundaries), and every read stays inside the stack allocation: the last pixel pair of the
// last row, at the largest frame offset, reads 16 bytes from its first sample.
inline bool pair_tile_fits(const void* stack, uint64_t alloc_bytes, uint32_t bps, uint32_t channels,
                           const OutlierK& k, uint64_t max_pos) {
    if ((uintptr_t)stack & 7u) return false;
    const uint64_t x_last = (uint64_t)((k.out_w - 1) / 2) * 2;
    const uint64_t first = max_pos + ((uint64_t)(k.out_h - 1) * k.stack_w + x_last) * channels;
    return first * bps + 16 <= alloc_bytes;
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
