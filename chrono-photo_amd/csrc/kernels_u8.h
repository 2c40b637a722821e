// kernels_u8.h -- tile shapes of the 8-bit register-resident outlier kernel and their launch.
//
// A shape (G, F, X, FS) puts G lanes on each group of X pixels, F frames of all positions in
// each lane (G * F >= N) and, for `--sample` composites, FS frames of the sample positions
// (G * FS >= S).
#pragma once
#include "k_outlier_u8.h"

namespace cp {

struct U8Shape {
    int G, F, X, FS, S;  // S: ranks from a sorting network (G = 1) instead of the byte search
};

// One pixel per lane for short stacks; four pixels per lane (one 16-byte load per frame)
// for tall ones; `--sample` shapes carry an FS-frame block of the sample positions; shapes
// 14-16 rank one lane's whole pixel with a sorting network; shape 17 holds a 64-frame pixel
// in one lane for the byte search (a candidate only: u8_shape_for meets shape 2 first).
constexpr int kNumU8Shapes = 18;
constexpr U8Shape kU8Shapes[kNumU8Shapes] = {
    {1, 16, 1, 0, 0},  {1, 32, 1, 0, 0},  {2, 32, 1, 0, 0},  {4, 32, 1, 0, 0}, {8, 32, 1, 0, 0},
    {16, 32, 1, 0, 0}, {16, 32, 4, 0, 0}, {32, 24, 4, 0, 0}, {32, 32, 4, 0, 0}, {32, 40, 4, 0, 0},
    {4, 32, 1, 8, 0},  {8, 32, 1, 8, 0},  {16, 32, 4, 8, 0}, {32, 40, 4, 4, 0}, {1, 16, 1, 0, 1},
    {1, 32, 1, 0, 1},  {1, 64, 1, 0, 1},  {1, 64, 1, 0, 0}};
constexpr uint32_t kU8MaxFrames = 32 * 40;
constexpr uint32_t kU8MaxSample = 128;

// The smallest shape that holds n frames (and s sample frames when sampled), or -1.
inline int u8_shape_for(uint32_t n, bool sampled, uint32_t s) {
    for (int i = 0; i < kNumU8Shapes; ++i) {
        const U8Shape& sh = kU8Shapes[i];
        if ((sh.FS > 0) != sampled || sh.S) continue;
        if ((uint32_t)(sh.G * sh.F) >= n && (!sampled || (uint32_t)(sh.G * sh.FS) >= s)) return i;
    }
    return -1;
}

template <int C, int G, int F, int X, int FS, bool SORT = false>
inline void launch_u8_shape(const OutlierK& k, hipStream_t s) {
    constexpr int P = X * 256 / G;
    const uint32_t tiles_x = (k.out_w + P - 1) / P;
    dim3 grid(tiles_x * k.out_h), block(256);
    if (k.th.absolute)
        hipLaunchKernelGGL((outlier_u8_kernel<C, G, F, X, FS, true, SORT>), grid, block, 0, s, k, tiles_x);
    else
        hipLaunchKernelGGL((outlier_u8_kernel<C, G, F, X, FS, false, SORT>), grid, block, 0, s, k, tiles_x);
}

template <int C>
inline bool launch_u8_c(const OutlierK& k, int shape, hipStream_t s) {
    switch (shape) {
    case 0: launch_u8_shape<C, 1, 16, 1, 0>(k, s); return true;
    case 1: launch_u8_shape<C, 1, 32, 1, 0>(k, s); return true;
    case 2: launch_u8_shape<C, 2, 32, 1, 0>(k, s); return true;
    case 3: launch_u8_shape<C, 4, 32, 1, 0>(k, s); return true;
    case 4: launch_u8_shape<C, 8, 32, 1, 0>(k, s); return true;
    case 5: launch_u8_shape<C, 16, 32, 1, 0>(k, s); return true;
    case 6: launch_u8_shape<C, 16, 32, 4, 0>(k, s); return true;
    case 7: launch_u8_shape<C, 32, 24, 4, 0>(k, s); return true;
    case 8: launch_u8_shape<C, 32, 32, 4, 0>(k, s); return true;
    case 9: launch_u8_shape<C, 32, 40, 4, 0>(k, s); return true;
    case 10: launch_u8_shape<C, 4, 32, 1, 8>(k, s); return true;
    case 11: launch_u8_shape<C, 8, 32, 1, 8>(k, s); return true;
    case 12: launch_u8_shape<C, 16, 32, 4, 8>(k, s); return true;
    case 13: launch_u8_shape<C, 32, 40, 4, 4>(k, s); return true;
    case 14: launch_u8_shape<C, 1, 16, 1, 0, true>(k, s); return true;
    case 15: launch_u8_shape<C, 1, 32, 1, 0, true>(k, s); return true;
    case 16: launch_u8_shape<C, 1, 64, 1, 0, true>(k, s); return true;
    case 17: launch_u8_shape<C, 1, 64, 1, 0>(k, s); return true;
    default: return false;
    }
}

// One translation unit per channel count: kernels_u8_c<C>.hip.
bool launch_u8_c1(const OutlierK& k, int shape, hipStream_t s);
bool launch_u8_c2(const OutlierK& k, int shape, hipStream_t s);
bool launch_u8_c3(const OutlierK& k, int shape, hipStream_t s);
bool launch_u8_c4(const OutlierK& k, int shape, hipStream_t s);

inline bool launch_outlier_u8(uint32_t channels, const OutlierK& k, int shape, hipStream_t s) {
    switch (channels) {
    case 1: return launch_u8_c1(k, shape, s);
    case 2: return launch_u8_c2(k, shape, s);
    case 3: return launch_u8_c3(k, shape, s);
    case 4: return launch_u8_c4(k, shape, s);
    default: return false;
    }
}

}  // namespace cp
