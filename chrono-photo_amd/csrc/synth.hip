// synth.hip -- libchrono_synth.so: seeded synthetic stacks on the host or in HBM.
// Not part of the compositing ABI; tests and bench.py use it to build workloads.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "synth.h"

#define CS_API __attribute__((visibility("default")))

extern "C" {

// Writes frames [first, first + n) restricted to rows [y0, y1) into host memory laid out
// [n][y1 - y0][width][channels].
CS_API int cs_generate_host(const cs_scene* s, uint32_t first, uint32_t n, uint32_t y0, uint32_t y1,
                     void* out) {
    if (!s || !out || y1 > s->height || y0 > y1) return -1;
    const uint32_t W = s->width, C = s->channels, rows = y1 - y0;
    for (uint32_t f = 0; f < n; ++f) {
        for (uint32_t y = y0; y < y1; ++y) {
            size_t base = (((size_t)f * rows + (y - y0)) * W) * C;
            for (uint32_t x = 0; x < W; ++x)
                for (uint32_t ch = 0; ch < C; ++ch) {
                    uint32_t v = cs_sample(s, first + f, x, y, ch);
                    if (s->bytes_per_sample == 1) ((uint8_t*)out)[base + (size_t)x * C + ch] = (uint8_t)v;
                    else ((uint16_t*)out)[base + (size_t)x * C + ch] = (uint16_t)v;
                }
        }
    }
    return 0;
}

CS_API void cs_shake_truth(const cs_scene* s, uint32_t n, int32_t* offsets) {
    for (uint32_t t = 0; t < n; ++t) cs_shake_offset(s, t, &offsets[2 * t], &offsets[2 * t + 1]);
}

}  // extern "C"

template <typename T>
__global__ void cs_generate_kernel(cs_scene s, uint32_t first, uint32_t n, T* __restrict__ out) {
    const uint64_t per_frame = (uint64_t)s.width * s.height * s.channels;
    const uint64_t total = per_frame * n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f = (uint32_t)(i / per_frame);
        uint64_t r = i - (uint64_t)f * per_frame;
        uint32_t ch = (uint32_t)(r % s.channels);
        uint64_t p = r / s.channels;
        uint32_t x = (uint32_t)(p % s.width);
        uint32_t y = (uint32_t)(p / s.width);
        out[i] = (T)cs_sample(&s, first + f, x, y, ch);
    }
}

extern "C" CS_API int cs_generate_device(const cs_scene* s, uint32_t first, uint32_t n, void* dev_out,
                                  void* stream) {
    if (!s || !dev_out) return -1;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t blocks = 256 * 32, threads = 256;
    if (s->bytes_per_sample == 1)
        hipLaunchKernelGGL(cs_generate_kernel<uint8_t>, dim3(blocks), dim3(threads), 0, st, *s, first,
                           n, (uint8_t*)dev_out);
    else
        hipLaunchKernelGGL(cs_generate_kernel<uint16_t>, dim3(blocks), dim3(threads), 0, st, *s,
                           first, n, (uint16_t*)dev_out);
    return hipGetLastError() == hipSuccess ? 0 : -6;
}
