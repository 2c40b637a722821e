// k_outlier_generic.h -- exact outlier composite for any stack height, one lane per pixel.
//
// This kernel covers every configuration the fast register-resident kernel does not
// (tall stacks, sample subsets, sums that are not exact in f32).  It reads samples straight
// from HBM/L2 and finds order statistics by an MSB-first bitwise search: for each needed
// rank r it keeps the largest prefix v with #{x < v} <= r, which is the r-th smallest
// value.  All ranks of one channel share each pass over the samples.
//
// Per-pixel semantics follow calc_pixel, src/chrono.rs:208-494: medians and inverse IQR
// on the sample set, weighted squared distance over all frames, background selection,
// outlier selection and blending with the same f32 operation order.
#pragma once
#include "device_common.h"

namespace cp {

// Ranks needed for the median (src/chrono.rs:580-591) and quartiles (:566-578) of n values.
struct RankPlan {
    uint32_t nr;
    uint32_t rank[6];
    // median: ranks m0, m1 (equal for odd n); q1: ranks a0,a1 + frac; q3: ranks b0,b1 + frac
    uint32_t m0, m1, a0, a1, b0, b1;
    float fa, fb;
    int qa_mode, qb_mode;  // 0: d[p1], 1: d[p1+1], 2: interpolate
};

__host__ __device__ inline void quantile_plan(uint32_t len, float q, uint32_t* r0, uint32_t* r1,
                                              float* frac_out, int* mode) {
    float pos = (float)(len + 1) * q;
    uint32_t p1 = (uint32_t)pos - 1u;
    float frac = pos - truncf(pos);
    *r0 = p1;
    *r1 = p1 + 1;
    *frac_out = frac;
    if (frac < 0.001f) { *mode = 0; *r1 = p1; }
    else if (frac > 0.999f) { *mode = 1; *r0 = p1 + 1; }
    else *mode = 2;
}

// The ranks and interpolation weights alone (no rank list).
__host__ __device__ inline RankPlan rank_plan_core(uint32_t len, bool absolute) {
    RankPlan p{};
    if ((len + 1) % 2 == 0) { p.m0 = p.m1 = (len + 1) / 2 - 1; }
    else { p.m1 = (len + 1) / 2; p.m0 = p.m1 - 1; }
    if (!absolute) {
        quantile_plan(len, 0.25f, &p.a0, &p.a1, &p.fa, &p.qa_mode);
        quantile_plan(len, 0.75f, &p.b0, &p.b1, &p.fb, &p.qb_mode);
    }
    return p;
}

__host__ __device__ inline RankPlan make_rank_plan(uint32_t len, bool absolute) {
    RankPlan p{};
    if ((len + 1) % 2 == 0) { p.m0 = p.m1 = (len + 1) / 2 - 1; }
    else { p.m1 = (len + 1) / 2; p.m0 = p.m1 - 1; }
    p.nr = 0;
    auto add = [&](uint32_t r) {
        for (uint32_t i = 0; i < p.nr; ++i) if (p.rank[i] == r) return;
        p.rank[p.nr++] = r;
    };
    add(p.m0);
    add(p.m1);
    if (!absolute) {
        quantile_plan(len, 0.25f, &p.a0, &p.a1, &p.fa, &p.qa_mode);
        quantile_plan(len, 0.75f, &p.b0, &p.b1, &p.fb, &p.qb_mode);
        add(p.a0); add(p.a1); add(p.b0); add(p.b1);
    }
    return p;
}

__device__ __forceinline__ uint32_t rank_value(const RankPlan& p, const uint32_t (&vals)[6], uint32_t r) {
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        if ((uint32_t)i < p.nr && p.rank[i] == r) v = vals[i];
    return v;
}

// median / quartile values from the selected order statistics (exact f32, same ops).
__device__ __forceinline__ void stats_from_ranks(const RankPlan& p, const uint32_t (&vals)[6],
                                                 bool absolute, float* median, float* iqr_inv) {
    float d0 = (float)rank_value(p, vals, p.m0), d1 = (float)rank_value(p, vals, p.m1);
    float med = p.m0 == p.m1 ? d0 : 0.5f * (d0 + d1);
    *median = med;
    if (absolute) { *iqr_inv = 0.0f; return; }
    auto quant = [&](uint32_t r0, uint32_t r1, float frac, int mode) -> float {
        float a = (float)rank_value(p, vals, r0), b = (float)rank_value(p, vals, r1);
        if (mode == 0) return a;
        if (mode == 1) return a;  // r0 was set to p1 + 1
        return (1.0f - frac) * a + frac * b;
    };
    float q1 = quant(p.a0, p.a1, p.fa, p.qa_mode);
    float q3 = quant(p.b0, p.b1, p.fb, p.qb_mode);
    float iqr = q3 - q1;
    if (iqr == 0.0f) iqr = 1.0f;
    *iqr_inv = 1.0f / iqr;
}

// stats_from_ranks with the values of ranks m0, m1, a0, a1, b0, b1 given directly.
__device__ __forceinline__ void stats_from_values(const RankPlan& p, uint32_t m0, uint32_t m1, uint32_t a0,
                                                  uint32_t a1, uint32_t b0, uint32_t b1, bool absolute,
                                                  float* median, float* iqr_inv) {
    float d0 = (float)m0, d1 = (float)m1;
    *median = p.m0 == p.m1 ? d0 : 0.5f * (d0 + d1);
    if (absolute) { *iqr_inv = 0.0f; return; }
    const float q1 = p.qa_mode == 2 ? (1.0f - p.fa) * (float)a0 + p.fa * (float)a1 : (float)a0;
    const float q3 = p.qb_mode == 2 ? (1.0f - p.fb) * (float)b0 + p.fb * (float)b1 : (float)b0;
    float iqr = q3 - q1;
    if (iqr == 0.0f) iqr = 1.0f;
    *iqr_inv = 1.0f / iqr;
}

template <int B, int C>
struct GenericPixel {
    const OutlierK& k;
    uint64_t pix;  // (y * stack_w + x) * C

    __device__ __forceinline__ uint32_t at(uint32_t t, int ch) const {
        return load_sample<B>(k.stack, k.pos_off[t] + pix + (uint64_t)ch);
    }
    __device__ __forceinline__ void sample(uint32_t t, uint32_t (&s)[4]) const {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) s[ch] = at(t, ch);
    }
    // src/chrono.rs:257-278
    __device__ __forceinline__ float dist_sq(uint32_t t, const float (&median)[4],
                                             const float (&iqr_inv)[4]) const {
        float d = 0.0f;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            float w = k.w[i];
            if (w == 0.0f) continue;
            float diff = median[i] - (float)at(t, i);
            float term;
            if (diff == 0.0f) term = 0.0f;
            else if (k.th.absolute) { float a = w * diff; term = signum_f(w) * (a * a); }
            else { float a = w * iqr_inv[i] * diff; term = signum_f(w) * (a * a); }
            d += term;
        }
        return d;
    }
    __device__ __forceinline__ bool in_sample(uint32_t t) const {
        return k.sample_mask == nullptr || ((k.sample_mask[t >> 5] >> (t & 31)) & 1u);
    }
};

template <int B, int C>
__global__ void __launch_bounds__(256) outlier_generic_kernel(OutlierK k) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= k.out_w || y >= k.out_h) return;
    constexpr uint32_t maxv = B == 1 ? 255u : 65535u;
    constexpr int BITS = 8 * B;
    GenericPixel<B, C> P{k, ((uint64_t)y * k.stack_w + x) * C};
    const uint32_t N = k.n;
    const uint32_t S = k.n_sample;
    const bool absolute = k.th.absolute != 0;
    const RankPlan plan = make_rank_plan(S, absolute);

    float median[4] = {0, 0, 0, 0}, iqr_inv[4] = {0, 0, 0, 0};
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        if (k.w[ch] == 0.0f) continue;
        uint32_t lo[6] = {0, 0, 0, 0, 0, 0};
        for (int bit = BITS - 1; bit >= 0; --bit) {
            uint32_t cnt[6] = {0, 0, 0, 0, 0, 0};
            uint32_t cand[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) cand[r] = lo[r] | (1u << bit);
            for (uint32_t t = 0; t < N; ++t) {
                if (!P.in_sample(t)) continue;
                uint32_t v = P.at(t, ch);
#pragma unroll
                for (int r = 0; r < 6; ++r) cnt[r] += v < cand[r] ? 1u : 0u;
            }
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if ((uint32_t)r < plan.nr && cnt[r] <= plan.rank[r]) lo[r] = cand[r];
        }
        stats_from_ranks(plan, lo, absolute, &median[ch], &iqr_inv[ch]);
    }

    const float threshold_sq = k.th.min * k.th.min;
    // Pass 1: classification summary (src/chrono.rs:257-288).
    uint32_t num = 0, max_index = 0, first_idx = 0, last_idx = 0, first_non = N;
    float max_d = 0.0f, first_d = 0.0f, last_d = 0.0f;
    for (uint32_t t = 0; t < N; ++t) {
        float d = P.dist_sq(t, median, iqr_inv);
        if (d >= threshold_sq) {
            if (num == 0) { first_idx = t; first_d = d; }
            last_idx = t; last_d = d;
            ++num;
            if (d > max_d) { max_d = d; max_index = t; }
        } else if (first_non == N) {
            first_non = t;
        }
    }

    uint32_t pixel[4] = {0, 0, 0, 0};
    bool warning = false;
    uint32_t bg_index = 0;
    const uint64_t pix_id = (uint64_t)y * k.out_w + x;
    // Background, src/chrono.rs:290-375.
    if (k.background == 2) {  // Average
        float mean[4] = {0, 0, 0, 0};
        for (uint32_t t = 0; t < N; ++t)
#pragma unroll
            for (int ch = 0; ch < C; ++ch) mean[ch] += (float)P.at(t, ch);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) mean[ch] /= (float)N;
        if (num == 1) {
#pragma unroll
            for (int ch = 0; ch < C; ++ch)
                pixel[ch] = sat_cast(roundf(mean[ch] * ((float)N / (float)(N - 1)) -
                                            (float)P.at(first_idx, ch) / (float)N), maxv);
        } else if (num > 1) {
            float osum[4] = {0, 0, 0, 0};
            for (uint32_t t = 0; t < N; ++t) {
                if (!(P.dist_sq(t, median, iqr_inv) >= threshold_sq)) continue;
#pragma unroll
                for (int ch = 0; ch < C; ++ch) osum[ch] += (float)P.at(t, ch);
            }
#pragma unroll
            for (int ch = 0; ch < C; ++ch)
                pixel[ch] = sat_cast(roundf(mean[ch] * ((float)N / (float)(N - num)) -
                                            osum[ch] / (float)N), maxv);
        } else {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(mean[ch]), maxv);
        }
    } else if (k.background == 3) {  // Median
#pragma unroll
        for (int ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(median[ch]), maxv);
    } else {
        if (k.background == 0) {  // First: first_excluded src/chrono.rs:505-530
            if (num == N) { bg_index = 0; warning = true; }
            else bg_index = num == 0 ? 0 : first_non;
        } else {  // Random: sample_excluded src/chrono.rs:532-555
            if (num == 0) {
                bg_index = rng_range(k.seed, k.pix_base + pix_id, 0, N);
            } else if (num == N) {
                bg_index = rng_range(k.seed, k.pix_base + pix_id, 0, N);
                warning = true;
            } else {
                // Position r of the swapped index array, traced back through the swaps
                // (a_j = j-th outlier, b_j = N - 1 - j) applied in order j = 0..num-1.
                uint32_t r = rng_range(k.seed, k.pix_base + pix_id, 0, N - num);
                uint32_t pos = r;
                uint32_t j = num;
                for (int64_t t = (int64_t)N - 1; t >= 0; --t) {
                    if (!(P.dist_sq((uint32_t)t, median, iqr_inv) >= threshold_sq)) continue;
                    --j;
                    uint32_t a = (uint32_t)t, b = N - 1 - j;
                    if (pos == a) pos = b;
                    else if (pos == b) pos = a;
                }
                bg_index = pos;
            }
        }
        uint32_t s[4];
        P.sample(bg_index, s);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) pixel[ch] = s[ch];
    }

    // Outlier selection and blending, src/chrono.rs:377-493.
    uint32_t blend_byte = 0;
    if (num > 0) {
        uint32_t s[4] = {0, 0, 0, 0};
        if (num == 1) {
            P.sample(first_idx, s);
            float fade = fade_pick(k.fade, (int32_t)first_idx, (int32_t)N, k.frame_offset);
            float blend = fade * blend_value(k.th, sqrtf(first_d));
            blend_into_px<C>(pixel, s, blend, maxv);
            blend_byte = sat_cast(roundf(blend * 255.0f), 255u);
        } else if (k.outlier == 4 || k.outlier == 5) {
            float pix_new[4] = {0, 0, 0, 0};
            float blend_inv = 1.0f;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) pix_new[ch] = (float)pixel[ch];
            const bool fwd = k.outlier == 4;
            for (uint32_t i = 0; i < N; ++i) {
                uint32_t t = fwd ? i : N - 1 - i;
                float d = P.dist_sq(t, median, iqr_inv);
                if (!(d >= threshold_sq)) continue;
                P.sample(t, s);
                float fade = fade_pick(k.fade, (int32_t)t, (int32_t)N, k.frame_offset);
                float blend = fade * blend_value(k.th, sqrtf(d));
                blend_into_f32<C>(pix_new, s, blend);
                blend_inv *= 1.0f - blend;
            }
#pragma unroll
            for (int ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(pix_new[ch]), maxv);
            blend_byte = sat_cast(roundf((1.0f - blend_inv) * 255.0f), 255u);
        } else {
            uint32_t sel = 0;
            float dist;
            if (k.outlier == 3) {  // Average (num > 1)
                float mean[4] = {0, 0, 0, 0};
                float mean_dist = 0.0f;
                for (uint32_t t = 0; t < N; ++t) {
                    float d = P.dist_sq(t, median, iqr_inv);
                    if (!(d >= threshold_sq)) continue;
#pragma unroll
                    for (int ch = 0; ch < C; ++ch) mean[ch] += (float)P.at(t, ch);
                    mean_dist += sqrtf(d);
                }
#pragma unroll
                for (int ch = 0; ch < C; ++ch) s[ch] = sat_cast(roundf(mean[ch] / (float)num), maxv);
                dist = mean_dist / (float)num;
            } else {
                float dsq;
                if (k.outlier == 0) { sel = first_idx; dsq = first_d; }
                else if (k.outlier == 1) { sel = last_idx; dsq = last_d; }
                else { sel = max_index; dsq = max_d; }
                P.sample(sel, s);
                dist = sqrtf(dsq);
            }
            float fade = fade_pick(k.fade, (int32_t)sel, (int32_t)N, k.frame_offset);
            float blend = fade * blend_value(k.th, dist);
            blend_into_px<C>(pixel, s, blend, maxv);
            blend_byte = sat_cast(roundf(blend * 255.0f), 255u);
        }
    }

    const uint64_t o = ((uint64_t)y * k.out_w + x) * C;
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        store_sample<B>(k.out_px, o + ch, pixel[ch]);
        if (k.out_mask) k.out_mask[o + ch] = ch < 3 ? (uint8_t)blend_byte : (uint8_t)255;
    }
    if (warning && k.warnings) atomicAdd(k.warnings, 1ull);
    if (k.stats) {
        PixelStats st;
#pragma unroll
        for (int i = 0; i < 4; ++i) { st.median[i] = median[i]; st.iqr_inv[i] = iqr_inv[i]; }
        st.num_outliers = num;
        st.max_index = max_index;
        st.max_dist_sq = max_d;
        st.bg_index = bg_index;
        st.warning = warning ? 1u : 0u;
        st.pad = 0;
        k.stats[(uint64_t)y * k.out_w + x] = st;
    }
}

}  // namespace cp
