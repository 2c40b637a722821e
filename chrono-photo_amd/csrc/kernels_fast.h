// kernels_fast.h -- tile-shape policy of the register-resident outlier kernel.
#pragma once
#include "k_outlier_fast.h"
#include "runtime.h"

namespace cp {

// Lanes per pixel (G) and frames per lane (F) for a stack of n frames: the smallest tile
// with G * F >= n.  Returns false when n exceeds the largest tile.
inline bool fast_tile_for(uint32_t n, int* G, int* F) {
    if (n <= 16) { *G = 1; *F = 16; return true; }
    if (n <= 32) { *G = 1; *F = 32; return true; }
    if (n <= 64) { *G = 2; *F = 32; return true; }
    if (n <= 128) { *G = 4; *F = 32; return true; }
    if (n <= 256) { *G = 8; *F = 32; return true; }
    if (n <= 512) { *G = 8; *F = 64; return true; }
    return false;
}

template <int B, int C, int G, int F>
inline void launch_fast_shape(const OutlierK& k, hipStream_t s) {
    constexpr int P = 256 / G;
    dim3 grid((k.out_w + P - 1) / P, k.out_h), block(256);
    if (k.th.absolute) hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, true>), grid, block, 0, s, k);
    else hipLaunchKernelGGL((outlier_fast_kernel<B, C, G, F, false>), grid, block, 0, s, k);
}

template <int B, int C>
inline bool launch_fast_c(const OutlierK& k, hipStream_t s) {
    int G, F;
    if (!fast_tile_for(k.n, &G, &F)) return false;
    if (G == 1 && F == 16) launch_fast_shape<B, C, 1, 16>(k, s);
    else if (G == 1 && F == 32) launch_fast_shape<B, C, 1, 32>(k, s);
    else if (G == 2) launch_fast_shape<B, C, 2, 32>(k, s);
    else if (G == 4 && F == 32) launch_fast_shape<B, C, 4, 32>(k, s);
    else if (G == 8 && F == 32) launch_fast_shape<B, C, 8, 32>(k, s);
    else launch_fast_shape<B, C, 8, 64>(k, s);
    return true;
}

bool launch_outlier_fast_u8(const cp_stack* st, const OutlierK& k, hipStream_t s);
bool launch_outlier_fast_u16(const cp_stack* st, const OutlierK& k, hipStream_t s);

}  // namespace cp
