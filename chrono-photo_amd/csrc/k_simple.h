// k_simple.h -- darker/lighter composite (SimpleProcessor, src/simple.rs:26-176).
//
// Per pixel, frames are visited in index order; value = sum_ch p * w[ch] in channel order
// (f32, no contraction); a strictly smaller (darker) or larger (lighter) value replaces the
// running extreme and blends the frame into the output with fade(t).  The output starts at
// zero and the extreme at +/-f32::MAX (src/simple.rs:77-85).
//
// Streaming kernel: each lane owns PX consecutive output pixels of one row (PX*C*B bytes, a
// multiple of 4 for every channel count with PX = 4 / B), so a wave reads one contiguous
// segment of a frame row; the stack is read exactly once.  A lane's segment is fetched as
// whole aligned dwords: all lanes of a row share the segment's offset within a dword (the row
// start and the frame offset are the only variable parts), so the shift that realigns the
// bytes is wave-uniform.  UF frames are loaded before the first of them is processed,
// keeping UF loads in flight per lane.
#pragma once
#include "device_common.h"

namespace cp {

// The NB bytes at address a as NB/4 dwords in d[0..NB/4) (d[NB/4] is scratch).  Only dwords
// that hold at least one byte of the segment are read.
template <int B, int NB>
__device__ __forceinline__ void load_segment(const uint8_t* a, uint32_t (&d)[NB / 4 + 1]) {
    const uint32_t s = (uint32_t)(uintptr_t)a & 3u;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(a - s);
    if (s == 0u) {
#pragma unroll
        for (int i = 0; i < NB / 4; ++i) d[i] = w[i];
        d[NB / 4] = 0u;
    } else {
#pragma unroll
        for (int i = 0; i <= NB / 4; ++i) d[i] = w[i];
#pragma unroll
        for (int i = 0; i < NB / 4; ++i) d[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], s);
    }
}

template <int B, int C, int PX>
__device__ __forceinline__ void unpack_segment(const uint32_t* d, uint32_t (&in)[PX][4]) {
#pragma unroll
    for (int e = 0; e < PX * C; ++e) {
        const uint32_t word = d[(e * B) / 4];
        const uint32_t sh = 8u * (uint32_t)((e * B) % 4);
        in[e / C][e % C] = B == 1 ? (word >> sh) & 0xFFu : (word >> sh) & 0xFFFFu;
    }
}

template <int B, int C, int PX>
__device__ __forceinline__ void simple_step(const SimpleK& k, bool darker, int32_t t, int32_t N,
                                            const uint32_t (&in)[PX][4], float (&ext)[PX], uint32_t (&out)[PX][4]) {
    constexpr uint32_t maxv = B == 1 ? 255u : 65535u;
    bool have_fade = false;  // fade(t) computed lazily, once per frame
    float fade = 1.0f;
#pragma unroll
    for (int p = 0; p < PX; ++p) {
        float value = 0.0f;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) value += (float)in[p][ch] * k.w[ch];
        const bool is_ext = darker ? (value < ext[p]) : (value > ext[p]);
        if (is_ext) {
            ext[p] = value;
            if (!have_fade) { fade = fade_pick(k.fade, t, N, k.frame_offset); have_fade = true; }
            const float f = fade;
            if (f > 0.0f) {
                if (f >= 1.0f) {
#pragma unroll
                    for (int ch = 0; ch < C; ++ch) out[p][ch] = in[p][ch];
                } else {
#pragma unroll
                    for (int ch = 0; ch < C; ++ch) {
                        const float aa = (float)out[p][ch];
                        out[p][ch] = sat_cast(roundf(aa + ((float)in[p][ch] - aa) * f), maxv);
                    }
                }
            }
        }
    }
}

template <int B, int C, int PX>
__global__ void __launch_bounds__(256) simple_kernel(SimpleK k) {
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * PX;
    const uint32_t y = blockIdx.y;
    if (x0 >= k.out_w || y >= k.out_h) return;
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
    constexpr int NB = PX * C * B;
    constexpr bool WHOLE = NB % 4 == 0;
    if (WHOLE && npx == (uint32_t)PX) {
        constexpr int UF = 8;
        int32_t t = 0;
        for (; t + UF <= N; t += UF) {
            uint32_t d[UF][NB / 4 + 1];
#pragma unroll
            for (int u = 0; u < UF; ++u) load_segment<B, NB>(k.stack + (k.pos_off[t + u] + pix) * B, d[u]);
#pragma unroll
            for (int u = 0; u < UF; ++u) {
                uint32_t in[PX][4];
                unpack_segment<B, C, PX>(d[u], in);
                simple_step<B, C, PX>(k, darker, t + u, N, in, ext, out);
            }
        }
        for (; t < N; ++t) {
            uint32_t d[NB / 4 + 1];
            load_segment<B, NB>(k.stack + (k.pos_off[t] + pix) * B, d);
            uint32_t in[PX][4];
            unpack_segment<B, C, PX>(d, in);
            simple_step<B, C, PX>(k, darker, t, N, in, ext, out);
        }
    } else {
        for (int32_t t = 0; t < N; ++t) {
            const uint64_t base = k.pos_off[t] + pix;
            uint32_t in[PX][4];
#pragma unroll
            for (int e = 0; e < PX * C; ++e)
                in[e / C][e % C] = (uint32_t)(e / C) < npx ? load_sample<B>(k.stack, base + e) : 0u;
            simple_step<B, C, PX>(k, darker, t, N, in, ext, out);
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
