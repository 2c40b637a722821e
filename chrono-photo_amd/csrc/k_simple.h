// k_simple.h -- darker/lighter composite (SimpleProcessor, src/simple.rs:26-176).
//
// Per pixel, frames are visited in index order; value = sum_ch p * w[ch] in channel order
// (f32, no contraction); a strictly smaller (darker) or larger (lighter) value replaces the
// running extreme and blends the frame into the output with fade(t).  The output starts at
// zero and the extreme at +/-f32::MAX (src/simple.rs:77-85).
//
// Streaming kernel: each lane owns PX consecutive output pixels of one row, so one frame
// row segment is read with wide coalesced loads when the row is aligned; the stack is read
// exactly once.
#pragma once
#include "device_common.h"

namespace cp {

template <int B, int C, int PX>
__global__ void __launch_bounds__(256) simple_kernel(SimpleK k) {
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * PX;
    const uint32_t y = blockIdx.y;
    if (x0 >= k.out_w || y >= k.out_h) return;
    constexpr uint32_t maxv = B == 1 ? 255u : 65535u;
    const uint32_t npx = min((uint32_t)PX, k.out_w - x0);
    const bool darker = k.darker != 0;
    float ext[PX];
    uint32_t out[PX][4];
#pragma unroll
    for (int p = 0; p < PX; ++p) {
        ext[p] = darker ? 3.40282347e+38f : -3.40282347e+38f;
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) out[p][ch] = 0;
    }
    const uint64_t pix = ((uint64_t)y * k.stack_w + x0) * C;
    const int32_t N = (int32_t)k.n;
    for (int32_t t = 0; t < N; ++t) {
        const uint64_t base = k.pos_off[t] + pix;
        uint32_t in[PX][4];
        if (npx == (uint32_t)PX) {
#pragma unroll
            for (int e = 0; e < PX * C; ++e) in[e / C][e % C] = load_sample<B>(k.stack, base + e);
        } else {
#pragma unroll
            for (int e = 0; e < PX * C; ++e)
                in[e / C][e % C] = (uint32_t)(e / C) < npx ? load_sample<B>(k.stack, base + e) : 0u;
        }
        bool have_fade = false;  // fade(t) computed lazily, once per frame
        float fade = 1.0f;
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            float value = 0.0f;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) value += (float)in[p][ch] * k.w[ch];
            bool is_ext = darker ? (value < ext[p]) : (value > ext[p]);
            if (is_ext) {
                ext[p] = value;
                if (!have_fade) { fade = fade_pick(k.fade, t, N, k.frame_offset); have_fade = true; }
                float f = fade;
                if (f > 0.0f) {
                    if (f >= 1.0f) {
#pragma unroll
                        for (int ch = 0; ch < C; ++ch) out[p][ch] = in[p][ch];
                    } else {
#pragma unroll
                        for (int ch = 0; ch < C; ++ch) {
                            float aa = (float)out[p][ch];
                            out[p][ch] = sat_cast(roundf(aa + ((float)in[p][ch] - aa) * f), maxv);
                        }
                    }
                }
            }
        }
    }
    const uint64_t o = ((uint64_t)y * k.out_w + x0) * C;
#pragma unroll
    for (int p = 0; p < PX; ++p) {
        if ((uint32_t)p >= npx) break;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) store_sample<B>(k.out_px, o + p * C + ch, out[p][ch]);
    }
}

}  // namespace cp
