// device_common.h -- f32 semantics shared by every compositing kernel.
//
// The reference computes in Rust f32 with no FMA contraction, correctly rounded division
// and square root, f32::round (half away from zero) and saturating `as u8` casts.  Every
// translation unit that includes this header is compiled with -ffp-contract=off and
// -fhip-fp32-correctly-rounded-divide-sqrt, so plain C++ float expressions below keep
// the reference's operation order and rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kparams.h"

namespace cp {

// Rust `f as u8` / `f as u16`: NaN -> 0, saturate to [0, maxv], truncate toward zero.
__device__ __forceinline__ uint32_t sat_cast(float f, uint32_t maxv) {
    if (!(f > 0.0f)) return 0u;
    if (f >= (float)maxv) return maxv;
    return (uint32_t)f;
}

// Rust f32::signum for finite non-NaN weights.
__device__ __forceinline__ float signum_f(float w) { return __builtin_signbitf(w) ? -1.0f : 1.0f; }

// Threshold::blend_value, src/options.rs:224-231.
__device__ __forceinline__ float blend_value(const ThresholdK& t, float dist) {
    if (dist <= t.min) return 0.0f;
    if (dist >= t.max) return 1.0f;
    return (dist - t.min) * t.scale;
}

// Fade::get src/options.rs:113-140 composed with the processors' fade()
// (src/chrono.rs:496-502, src/simple.rs:170-176).
__device__ __forceinline__ float fade_pick(const FadeK& f, int32_t frame, int32_t total,
                                           int32_t offset) {
    if (f.is_none) return 1.0f;
    int32_t q = f.absolute ? offset + frame : total - frame - 1;
    int32_t i = q - f.offset;
    int32_t len = (int32_t)f.len;
    if (i >= 0 && i < len) return f.values[i];
    if (!f.repeat) return i < 0 ? f.values[0] : f.values[len - 1];
    i %= len;
    if (i < 0) i += len;
    return f.values[i];
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Seeded replacement for the reference's ThreadRng::gen_range(0, n).
__device__ __forceinline__ uint32_t rng_range(uint64_t seed, uint64_t pixel, uint32_t draw,
                                              uint32_t n) {
    uint64_t h = splitmix64(seed ^ splitmix64(pixel * 4ull + (uint64_t)draw));
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

template <int B>
__device__ __forceinline__ uint32_t load_sample(const uint8_t* __restrict__ base, uint64_t idx) {
    if constexpr (B == 1) return base[idx];
    else return reinterpret_cast<const uint16_t*>(base)[idx];
}

template <int B>
__device__ __forceinline__ void store_sample(void* __restrict__ out, uint64_t idx, uint32_t v) {
    if constexpr (B == 1) reinterpret_cast<uint8_t*>(out)[idx] = (uint8_t)v;
    else reinterpret_cast<uint16_t*>(out)[idx] = (uint16_t)v;
}

// color::blend_into_u8 src/color.rs:4-16 on C channels (u8 or u16 saturation).
template <int C>
__device__ __forceinline__ void blend_into_px(uint32_t (&a)[4], const uint32_t (&b)[4], float blend,
                                              uint32_t maxv) {
    if (blend <= 0.0f) return;
    if (blend >= 1.0f) {
#pragma unroll
        for (int i = 0; i < C; ++i) a[i] = b[i];
        return;
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        float aa = (float)a[i];
        a[i] = sat_cast(roundf(aa + ((float)b[i] - aa) * blend), maxv);
    }
}

// color::blend_into_f32_u8 src/color.rs:32-44.
template <int C>
__device__ __forceinline__ void blend_into_f32(float (&a)[4], const uint32_t (&b)[4], float blend) {
    if (blend <= 0.0f) return;
    if (blend >= 1.0f) {
#pragma unroll
        for (int i = 0; i < C; ++i) a[i] = (float)b[i];
        return;
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        float aa = a[i];
        a[i] = aa + ((float)b[i] - aa) * blend;
    }
}

}  // namespace cp
