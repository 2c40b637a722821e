// k_outlier_fast.h -- register-resident outlier composite (the hot path).
//
// Layout.  A workgroup of 256 lanes covers P = X*256/G consecutive output pixels of one
// row.  The G lanes of a pixel group (adjacent lanes, one DPP quad or half-row) each hold a
// contiguous block of F frames of the group's X pixels x C channels in VGPRs, two samples
// per VGPR (16-bit halves, frames 2j and 2j+1), so the stack tile is read from HBM exactly
// once and stays on chip for the order-statistic pass and the classification pass.  With
// X = 2 (16-bit stacks) each frame of a lane's block is one 16-byte load, which is what lets
// the lanes' loads run at HBM speed; one pixel per lane needs two or three narrow loads per
// frame and caps the read rate well below it.  Frame t's element offset (stack slot + crop
// origin) comes from a per-workgroup LDS table padded to G*F entries.
//
// Order statistics (src/chrono.rs:228-255 sorts each weighted channel; any exact selection
// gives the same values).  The r-th smallest value is found MSB-first: keep the largest
// prefix v with #{y < v} <= r.  Counts come from the sum of absolute differences:
//     sum_y |y - v| - sum_y |y - (v-1)| = 2 #{y < v} - n
// so one search step costs two v_sad_u8 per four byte-samples plus a G-lane DPP reduction.
// All channels (and, for relative thresholds, all ranks) advance through the same step, so
// the independent chains interleave.  8-bit stacks need one 8-step level, which their medians
// first try as 4 steps in a 16-value window around an estimate and their quartiles as 4 steps
// next to the median; 16-bit stacks run one level on the high bytes and one on the low bytes
// clamped into the winning high-byte bucket (values below map to 0, above to 255, which
// leaves every count #{y < v}, v in [1,255], exact); their medians first try a single 7-step
// level inside a 128-wide window around an estimate, then the window next to it, and fall
// back to the two levels when the result lands on a window's edge.  The neighbour order statistic of an even-sized median
// (and the quartile interpolation partners) is either the same value (when #{x <= v} covers
// it) or min{x > v} (packed u16 min).  Frames beyond N are padded with the maximum sample
// value, which never changes the ranks below N.
//
// Classification (src/chrono.rs:257-288) runs over the lane's frames with the reference's
// f32 operation order; outliers become bits of a per-lane mask.  Each lane then leaves its
// mask and its largest outlier distance in LDS, and after one barrier the lanes finish one
// pixel each (background and outlier selection / blending, src/chrono.rs:290-493), so the
// per-pixel tail runs on full waves instead of one lane in G.  The finish re-reads from L2
// only the few samples an order-dependent blend needs.
#pragma once
#include <type_traits>

#include "device_common.h"
#include "k_outlier_generic.h"

namespace cp {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 as_us2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t as_u32(us2 v) { return __builtin_bit_cast(uint32_t, v); }

// Sum / min / max over the G adjacent lanes of a pixel group (G in {1,2,4,8,16,32}); the DPP
// moves fold into v_add/v_min/v_max_u32_dpp.  Every lane of the group receives the result:
// xor 1 and xor 2 within quads, then the half-row mirror (lane i <-> 7 - i) joins the two
// quads of each half-row and the row mirror (lane i <-> 15 - i) the two half-rows; for 32
// lanes v_permlane16_swap of the row results gives each lane of a row pair both rows
// (its first result holds the even row of the pair, its second the odd row).
template <int G>
__device__ __forceinline__ int32_t grp_sum(int32_t v) {
    if constexpr (G >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);    // xor 1
    if constexpr (G >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);    // xor 2
    if constexpr (G >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);   // half-row mirror
    if constexpr (G >= 16) v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true);  // row mirror
    if constexpr (G >= 32) {
        const auto r = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
        v = (int32_t)(r[0] + r[1]);
    }
    return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_min_u32(uint32_t v) {
    if constexpr (G >= 2) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0xB1, 0xF, 0xF, false));
    if constexpr (G >= 4) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x4E, 0xF, 0xF, false));
    if constexpr (G >= 8) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x141, 0xF, 0xF, false));
    if constexpr (G >= 16) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x140, 0xF, 0xF, false));
    if constexpr (G >= 32) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = min((uint32_t)r[0], (uint32_t)r[1]);
    }
    return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_max_u32(uint32_t v) {
    if constexpr (G >= 2) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    if constexpr (G >= 4) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    if constexpr (G >= 8) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    if constexpr (G >= 16) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    if constexpr (G >= 32) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = max((uint32_t)r[0], (uint32_t)r[1]);
    }
    return v;
}

// For every channel c and rank k[r]: the largest v in [0, 2^BITS - 1] with #{bytes < v} <=
// k[r] over the group's ND dwords per lane of y[c] (bytes below 2^BITS when BITS < 8).
// Candidates are kept replicated in all four bytes, so each step is two v_sad_u8 per dword, a
// G-lane reduction and a select per (channel, rank); the C*R chains are independent and
// interleave.  With CNT, cnt[c][r] receives #{bytes <= v}: v + 1 is the last rejected
// candidate (v = prefix 0 1..1 and the step at that 0 bit tested prefix 1 0..0), whose count
// the search already computed; no rejection means v = 2^BITS - 1 and every byte counts.
template <int G, int ND, int C, int R, bool CNT, int BITS = 8>
__device__ __forceinline__ void byte_select_multi(const uint32_t (&y)[C][ND], const uint32_t (&k)[R],
                                                  int32_t ntot, uint32_t (&out)[C][R], uint32_t (&cnt)[C][R]) {
    static_assert(BITS >= 1 && BITS <= 8, "byte search width");
    uint32_t lo[C][R];
    int32_t drej[C][R], thr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) thr[r] = 2 * (int32_t)k[r] - ntot;
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            lo[c][r] = 0;
            drej[c][r] = ntot;
        }
#pragma unroll 1
    for (int bit = BITS - 1; bit >= 0; --bit) {
        const uint32_t inc = 0x01010101u << bit;
        const uint32_t below = inc - 0x01010101u;  // the lower bits of every byte
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                // lo has no bits at or below `bit`, so V - 1 (per byte) is lo | below
                const uint32_t V = lo[c][r] | inc, Vm = lo[c][r] | below;
                uint32_t a = 0, b = 0;
#pragma unroll
                for (int i = 0; i < ND; ++i) {
                    a = __builtin_amdgcn_sad_u8(y[c][i], V, a);
                    b = __builtin_amdgcn_sad_u8(y[c][i], Vm, b);
                }
                const int32_t d = grp_sum<G>((int32_t)(a - b));  // 2 #{y < v} - ntot
                if (d <= thr[r]) lo[c][r] = V;
                else if constexpr (CNT) drej[c][r] = d;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            out[c][r] = lo[c][r] & 0xFFu;
            if constexpr (CNT) cnt[c][r] = (uint32_t)(drej[c][r] + ntot) >> 1;
        }
}

// 2 #{y < v} - ntot over the group's packed bytes, for v in 1..255.
template <int G, int ND>
__device__ __forceinline__ int32_t count_below(const uint32_t (&y)[ND], uint32_t v) {
    const uint32_t V = v * 0x01010101u, Vm = (v - 1u) * 0x01010101u;
    uint32_t a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        a = __builtin_amdgcn_sad_u8(y[i], V, a);
        b = __builtin_amdgcn_sad_u8(y[i], Vm, b);
    }
    return grp_sum<G>((int32_t)(a - b));
}

// us2 median of three (per 16-bit half).
__device__ __forceinline__ us2 med3_us2(us2 a, us2 b, us2 c) {
    const us2 lo = __builtin_elementwise_min(a, b), hi = __builtin_elementwise_max(a, b);
    return __builtin_elementwise_max(lo, __builtin_elementwise_min(hi, c));
}

// Location estimate of one channel's packed bytes over the group, which only places a search
// window (exactness never depends on it).  A lane with its first 12 frames real (ok) offers
// the mean of four medians of three of them; the group takes the mean of the offers without
// the smallest and the largest (all of them with fewer than three), nv = the number of
// offering lanes (nonzero).
template <int G, int ND>
__device__ __forceinline__ uint32_t byte_estimate(const uint32_t (&y)[ND], bool ok, uint32_t nv) {
    static_assert(ND >= 3, "estimate reads three dwords");
    us2 a[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {  // frames 4i, 4i+2 and 4i+1, 4i+3 widened to 16-bit halves
        a[2 * i] = as_us2(__builtin_amdgcn_perm(0u, y[i], 0x0C020C00u));
        a[2 * i + 1] = as_us2(__builtin_amdgcn_perm(0u, y[i], 0x0C030C01u));
    }
    const us2 m = med3_us2(a[0], a[1], a[2]), n = med3_us2(a[3], a[4], a[5]);
    const uint32_t e = ((uint32_t)m.x + (uint32_t)m.y + (uint32_t)n.x + (uint32_t)n.y) >> 2;
    if constexpr (G == 1) return e;
    const uint32_t s = (uint32_t)grp_sum<G>((int32_t)(ok ? e : 0u));
    uint32_t num = s, den = nv;
    if constexpr (G >= 4) {
        if (nv >= 3u) {
            num = s - grp_min_u32<G>(ok ? e : 0xFFFFu) - grp_max_u32<G>(ok ? e : 0u);
            den = nv - 2u;
        }
    }
    return den == 1u ? num : (uint32_t)((float)num * __builtin_amdgcn_rcpf((float)den));
}

// The k-th smallest of every channel's packed bytes (out) and #{bytes <= it} (cnt), as
// byte_select_multi's single rank: a 4-step search in the 16 values [lo, lo + 15] around the
// estimate, lo = clamp(estimate - 7, 0, 240).  The result v = lo + j is exact when 1 <= j <= 14:
// #{y < v} <= k was counted when j's last 1 bit was set and #{y < v + 1} > k when that
// candidate was rejected; j = 0 at lo = 0 and j = 15 at lo = 240 are exact too (the value is 0
// or 255).  A channel whose result lands on another edge runs the 8-step search; so does every
// channel when no lane of the group has 12 real frames (real: the lane's frames below the
// stack's count).
template <int G, int ND, int C>
__device__ __forceinline__ void byte_select_median(const uint32_t (&y)[C][ND], uint32_t k, int32_t ntot,
                                                   uint32_t real, uint32_t (&out)[C], uint32_t (&cnt)[C]) {
    const uint32_t kk[1] = {k};
    const bool ok = real >= 12u;
    const uint32_t nv = G == 1 ? (ok ? 1u : 0u) : (uint32_t)grp_sum<G>(ok ? 1 : 0);
    if (ND < 3 || nv == 0u) {  // uniform over the wave: the same frame count in every group
        uint32_t o[C][1], n[C][1];
        byte_select_multi<G, ND, C, 1, true>(y, kk, ntot, o, n);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            out[c] = o[c][0];
            cnt[c] = n[c][0];
        }
        return;
    }
    if constexpr (ND >= 3) {
        const int32_t thr = 2 * (int32_t)k - ntot;
        uint32_t lo[C], j[C];
        int32_t drej[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const uint32_t est = byte_estimate<G, ND>(y[c], ok, nv);
            lo[c] = min(est >= 7u ? est - 7u : 0u, 240u);
            j[c] = 0;
            drej[c] = ntot;
        }
#pragma unroll 1
        for (int bit = 3; bit >= 0; --bit) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int32_t d = count_below<G, ND>(y[c], lo[c] + j[c] + (1u << bit));
                if (d <= thr) j[c] += 1u << bit;
                else drej[c] = d;
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if ((j[c] >= 1u || lo[c] == 0u) && (j[c] <= 14u || lo[c] == 240u)) {
                out[c] = lo[c] + j[c];
                cnt[c] = (uint32_t)(drej[c] + ntot) >> 1;
            } else {
                uint32_t one[1][ND], o[1][1], n[1][1];
#pragma unroll
                for (int i = 0; i < ND; ++i) one[0][i] = y[c][i];
                byte_select_multi<G, ND, 1, 1, true>(one, kk, ntot, o, n);
                out[c] = o[0][0];
                cnt[c] = n[0][0];
            }
        }
    }
}

// The median and both quartiles of every channel's packed bytes: k = {median rank, first quartile
// rank <= it, third quartile rank >= it}; out and cnt as byte_select_multi's.  The median comes
// from byte_select_median (real as there).  The first quartile then lies in the 16 values
// [lo, lo + 15] with lo = max(0, median - 15) when #{y < lo} <= its rank (the window's top,
// median + 1 or above, already counts more than the median's rank), and the third quartile in
// [lo, lo + 15] with lo = min(median, 240) when #{y < lo + 16} > its rank (#{y < lo} is at most
// the median's rank): each takes 4 steps instead of 8.  When the two far-end counts show a
// quartile outside its window for any lane of the wave, the wave runs the full search for both
// quartiles instead; either way the values are the exact ranks.
template <int G, int ND, int C>
__device__ __forceinline__ void byte_select_quartiles(const uint32_t (&y)[C][ND], const uint32_t (&k)[3], int32_t ntot,
                                                      uint32_t real, uint32_t (&out)[C][3], uint32_t (&cnt)[C][3]) {
    {
        uint32_t o[C], n[C];
        byte_select_median<G, ND, C>(y, k[0], ntot, real, o, n);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            out[c][0] = o[c];
            cnt[c][0] = n[c];
        }
    }
    const int32_t thrA = 2 * (int32_t)k[1] - ntot, thrB = 2 * (int32_t)k[2] - ntot;
    uint32_t loA[C], loB[C], cA[C], cB[C];
    int32_t drA[C], drB[C];
    bool outside = k[1] > k[0] || k[2] < k[0];  // grid-uniform
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const uint32_t m = out[c][0];
        loA[c] = m >= 15u ? m - 15u : 0u;
        loB[c] = min(m, 240u);
        const int32_t dA = count_below<G, ND>(y[c], max(loA[c], 1u));
        const int32_t dB = count_below<G, ND>(y[c], min(loB[c] + 16u, 255u));
        outside = outside || (loA[c] > 0u && dA > thrA) || (loB[c] < 240u && dB <= thrB);
        drA[c] = 2 * (int32_t)cnt[c][0] - ntot;       // #{y < median + 1}
        drB[c] = loB[c] < 240u ? dB : ntot;            // #{y < lo + 16}
        cA[c] = cB[c] = 0;
    }
    if (__any(outside)) {
        const uint32_t kq[2] = {k[1], k[2]};
        uint32_t o[C][2], n[C][2];
        byte_select_multi<G, ND, C, 2, true>(y, kq, ntot, o, n);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            out[c][1] = o[c][0];
            cnt[c][1] = n[c][0];
            out[c][2] = o[c][1];
            cnt[c][2] = n[c][1];
        }
        return;
    }
#pragma unroll 1
    for (int bit = 3; bit >= 0; --bit) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const uint32_t VA = loA[c] + cA[c] + (1u << bit), VB = loB[c] + cB[c] + (1u << bit);
            const int32_t dA = count_below<G, ND>(y[c], VA), dB = count_below<G, ND>(y[c], VB);
            if (dA <= thrA) cA[c] += 1u << bit;
            else drA[c] = dA;
            if (dB <= thrB) cB[c] += 1u << bit;
            else drB[c] = dB;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        out[c][1] = loA[c] + cA[c];
        cnt[c][1] = (uint32_t)(drA[c] + ntot) >> 1;
        out[c][2] = loB[c] + cB[c];
        cnt[c][2] = (uint32_t)(drB[c] + ntot) >> 1;
    }
}

// Low bytes of the 16-bit samples (8-bit stacks keep one sample per 16-bit half).
template <int NW>
__device__ __forceinline__ void pack_low_bytes(const uint32_t (&r)[NW], uint32_t (&y)[NW / 2]) {
#pragma unroll
    for (int i = 0; i < NW / 2; ++i) y[i] = __builtin_amdgcn_perm(r[2 * i + 1], r[2 * i], 0x06040200u);
}

// clamp(x - base, 0, TOP) of every 16-bit sample (TOP <= 255), packed four to a dword.
template <int NW, uint32_t TOP = 255>
__device__ __forceinline__ void pack_window(const uint32_t (&r)[NW], uint32_t base, uint32_t (&y)[NW / 2]) {
    const us2 b2 = {(unsigned short)base, (unsigned short)base};
    const us2 ctop = {(unsigned short)TOP, (unsigned short)TOP};
#pragma unroll
    for (int i = 0; i < NW / 2; ++i) {
        us2 a = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i]), b2), ctop);
        us2 b = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i + 1]), b2), ctop);
        y[i] = __builtin_amdgcn_perm(as_u32(b), as_u32(a), 0x06040200u);
    }
}

// Exact r-th smallest of every channel of a 16-bit stack by two byte levels.
template <int G, int NW, int C>
__device__ __forceinline__ void select_rank_u16_multi(const uint32_t (&raw)[C][NW], uint32_t k, int32_t ntot,
                                                      uint32_t (&v)[C]) {
    constexpr int ND = NW / 2;
    uint32_t y[C][ND];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int i = 0; i < ND; ++i) y[c][i] = __builtin_amdgcn_perm(raw[c][2 * i + 1], raw[c][2 * i], 0x07050301u);
    const uint32_t kk[1] = {k};
    uint32_t H[C][1], L[C][1], unused[C][1];
    byte_select_multi<G, ND, C, 1, false>(y, kk, ntot, H, unused);
#pragma unroll
    for (int c = 0; c < C; ++c) pack_window<NW>(raw[c], H[c][0] << 8, y[c]);
    byte_select_multi<G, ND, C, 1, false>(y, kk, ntot, L, unused);
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = (H[c][0] << 8) | L[c][0];
}

// #{x <= v} over the group (v below the maximum sample value).
template <int G, int NW>
__device__ __forceinline__ uint32_t count_le(const uint32_t (&r)[NW], uint32_t v, int32_t ntot) {
    const uint32_t V0 = v | (v << 16), V1 = (v + 1) | ((v + 1) << 16);
    uint32_t a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        a = __builtin_amdgcn_sad_u16(r[i], V1, a);
        b = __builtin_amdgcn_sad_u16(r[i], V0, b);
    }
    const int32_t d = grp_sum<G>((int32_t)(a - b));  // 2 #{x <= v} - n
    return (uint32_t)(ntot + d) >> 1;
}

// min{x > v} over the group (exists by the caller's count).
template <int G, int NW>
__device__ __forceinline__ uint32_t min_greater(const uint32_t (&r)[NW], uint32_t v) {
    const us2 V = {(unsigned short)(v + 1), (unsigned short)(v + 1)};
    us2 m = {0xFFFF, 0xFFFF};
#pragma unroll
    for (int i = 0; i < NW; ++i) m = __builtin_elementwise_min(m, as_us2(r[i]) - V);
    uint32_t mm = min((uint32_t)m.x, (uint32_t)m.y);
    return grp_min_u32<G>(mm) + v + 1;
}

// Location estimate of a 16-bit channel: median-of-3 "remedian" over the lane's samples
// (both halves of each dword at once), then a trimmed mean over the group's lanes.  It only
// places the search window; exactness never depends on it.
template <int G, int NW>
__device__ __forceinline__ uint32_t group_estimate(const uint32_t (&r)[NW]) {
    constexpr int L0 = NW / 3, L1 = L0 / 3, L2 = L1 / 3;
    static_assert(L0 >= 1, "estimate needs at least three dwords per lane");
    us2 a[L0];
#pragma unroll
    for (int i = 0; i < L0; ++i) a[i] = med3_us2(as_us2(r[3 * i]), as_us2(r[3 * i + 1]), as_us2(r[3 * i + 2]));
    if constexpr (L1 >= 1) {
#pragma unroll
        for (int i = 0; i < L1; ++i) a[i] = med3_us2(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    }
    if constexpr (L2 >= 1) {
#pragma unroll
        for (int i = 0; i < L2; ++i) a[i] = med3_us2(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    }
    const uint32_t e = ((uint32_t)a[0].x + (uint32_t)a[0].y) >> 1;
    if constexpr (G == 1) return e;
    const uint32_t s = (uint32_t)grp_sum<G>((int32_t)e);
    if constexpr (G == 2) return s >> 1;
    return (s - grp_min_u32<G>(e) - grp_max_u32<G>(e)) / (uint32_t)(G - 2);
}

// Neighbour rank r1 = r0 + 1 from v0 and cle = #{x <= v0}.
template <int G, int NW>
__device__ __forceinline__ uint32_t neighbour(const uint32_t (&r)[NW], uint32_t v0, uint32_t cle, uint32_t r1) {
    return cle >= r1 + 1 ? v0 : min_greater<G, NW>(r, v0);
}

// Values at ranks (r0, r1), r1 in {r0, r0 + 1}, of every channel.
//  u8:  one byte level (WINDOW: in the 16-value window of byte_select_median, real = the lane
//       block's frames below the stack's count); the neighbour is v0 or min{x > v0}.
//  u16, WINDOW (medians of absolute thresholds): a 7-step byte search inside the window
//       [lo, lo + 127] around the estimate; a channel's result is exact when it ends strictly
//       inside (1..126).  Otherwise the rank lies below lo + 1 (result 0) or above lo + 126
//       (result 127), and that channel searches the window next to it on that side, then, if
//       that one misses too, runs the exact two-level search.
//  u16 otherwise: the exact two-level search.
// Branches depend on group-uniform values only.
template <int B, int G, int NW, int C, bool WINDOW>
__device__ __forceinline__ void rank_pairs(const uint32_t (&raw)[C][NW], uint32_t r0, uint32_t r1, int32_t ntot,
                                           uint32_t real, uint32_t (&v0)[C], uint32_t (&v1)[C]) {
    constexpr int ND = NW / 2;
    constexpr uint32_t MAXV = B == 1 ? 255u : 65535u;
    const uint32_t kk[1] = {r0};
    uint32_t sel[C][1], cle[C][1];
    if constexpr (B == 1) {
        uint32_t y[C][ND];
#pragma unroll
        for (int c = 0; c < C; ++c) pack_low_bytes<NW>(raw[c], y[c]);
        uint32_t s1[C], n1[C];
        if constexpr (WINDOW) {
            byte_select_median<G, ND, C>(y, r0, ntot, real, s1, n1);
        } else {
            byte_select_multi<G, ND, C, 1, true>(y, kk, ntot, sel, cle);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                s1[c] = sel[c][0];
                n1[c] = cle[c][0];
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            v0[c] = s1[c];
            v1[c] = r1 == r0 ? v0[c] : neighbour<G, NW>(raw[c], v0[c], n1[c], r1);
        }
    } else if constexpr (WINDOW) {
        constexpr uint32_t TOP = 127u, LO_MAX = 65535u - TOP;
        uint32_t y[C][ND], lo[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const uint32_t est = group_estimate<G, NW>(raw[c]);
            lo[c] = min(est >= 63u ? est - 63u : 0u, LO_MAX);
            pack_window<NW, TOP>(raw[c], lo[c], y[c]);
        }
        byte_select_multi<G, ND, C, 1, true, 7>(y, kk, ntot, sel, cle);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uint32_t ys = sel[c][0], base = lo[c];
            uint32_t le = cle[c][0];  // #{x <= v0} when v0 lies inside the window
            if (ys < 1u || ys > TOP - 1u) {
                base = ys == 0u ? (base >= TOP - 1u ? base - (TOP - 1u) : 0u) : min(base + (TOP - 1u), LO_MAX);
                uint32_t y1[1][ND], s1[1][1], c1[1][1];
                pack_window<NW, TOP>(raw[c], base, y1[0]);
                byte_select_multi<G, ND, 1, 1, true, 7>(y1, kk, ntot, s1, c1);
                ys = s1[0][0];
                le = c1[0][0];
            }
            if (ys >= 1u && ys <= TOP - 1u) {
                v0[c] = base + ys;
            } else {
                uint32_t one[1][NW];
#pragma unroll
                for (int i = 0; i < NW; ++i) one[0][i] = raw[c][i];
                uint32_t v[1];
                select_rank_u16_multi<G, NW, 1>(one, r0, ntot, v);
                v0[c] = v[0];
                if (r1 != r0) le = v0[c] >= MAXV ? (uint32_t)ntot : count_le<G, NW>(raw[c], v0[c], ntot);
            }
            v1[c] = r1 == r0 ? v0[c] : neighbour<G, NW>(raw[c], v0[c], le, r1);
        }
    } else {
        select_rank_u16_multi<G, NW, C>(raw, r0, ntot, v0);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (r1 == r0) { v1[c] = v0[c]; continue; }
            const uint32_t cle = v0[c] >= MAXV ? (uint32_t)ntot : count_le<G, NW>(raw[c], v0[c], ntot);
            v1[c] = neighbour<G, NW>(raw[c], v0[c], cle, r1);
        }
    }
}

// The three rank pairs of a relative threshold (median, first and third quartile) of an
// 8-bit stack in one byte search: the bytes are packed once and the 3*C chains interleave.
template <int G, int NW, int C>
__device__ __forceinline__ void rank_triples_u8(const uint32_t (&raw)[C][NW], const RankPlan& plan, int32_t ntot,
                                                uint32_t real, uint32_t (&m0)[C], uint32_t (&m1)[C], uint32_t (&a0)[C],
                                                uint32_t (&a1)[C], uint32_t (&b0)[C], uint32_t (&b1)[C]) {
    constexpr int ND = NW / 2;
    const uint32_t kk[3] = {plan.m0, plan.a0, plan.b0};
    uint32_t y[C][ND], sel[C][3], cle[C][3];
#pragma unroll
    for (int c = 0; c < C; ++c) pack_low_bytes<NW>(raw[c], y[c]);
    byte_select_quartiles<G, ND, C>(y, kk, ntot, real, sel, cle);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        m0[c] = sel[c][0];
        a0[c] = sel[c][1];
        b0[c] = sel[c][2];
        m1[c] = plan.m1 == plan.m0 ? m0[c] : neighbour<G, NW>(raw[c], m0[c], cle[c][0], plan.m1);
        a1[c] = plan.a1 == plan.a0 ? a0[c] : neighbour<G, NW>(raw[c], a0[c], cle[c][1], plan.a1);
        b1[c] = plan.b1 == plan.b0 ? b0[c] : neighbour<G, NW>(raw[c], b0[c], cle[c][2], plan.b1);
    }
}

// src/chrono.rs:257-278 on one frame's C samples (f32, reference operation order).
template <int C>
__device__ __forceinline__ float frame_dist(const uint32_t (&s)[4], const float (&med)[4],
                                            const float (&wi)[4], const float (&sg)[4]) {
    float d = 0.0f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const float diff = med[c] - (float)s[c];
        const float a = wi[c] * diff;
        d += sg[c] * (a * a);
    }
    return d;
}

// Per-tile LDS: the frame-offset table, every lane's classification summary and every
// pixel's order statistics.
// Layouts keep every access of a wave in distinct LDS banks: the offset table leaves two
// spare entries after each lane block, so the G lanes of a pixel (which read entries
// g*F + j at once) start in different banks; per-lane summaries sit at slot(g, p), which
// rotates each frame block's row (by 4g for one pixel per lane; for two pixels per lane by
// an odd/even pair of offsets that spreads the even and odd pixels) so that the finishing
// lanes (fixed g, consecutive p) and the writing lanes (the pixels of a wave x G blocks)
// both spread over the banks; per-pixel values are stored channel-major.
template <int C, int G, int F, int X = 1>
struct TileShared {
    static constexpr int P = X * 256 / G;      // pixels per tile
    static constexpr int NS = G * P;           // lane summary slots (X per lane)
    static constexpr int MW = (F + 31) / 32;   // mask words per lane and pixel
    static constexpr int OS = F + 2;           // offset-table stride of a lane block
    uint64_t off[G * OS];      // frame byte offsets at oidx(t); positions >= N repeat N-1
    // Two pixels per lane: with r the tile row's first sample address modulo 4 (shared by
    // every pixel pair of the row, which start 12 or 16 bytes apart), aoff = (off + r) & ~3
    // and sft = (off + r) & 3, so the lane whose pair starts at base loads 16 bytes at
    // (base & ~3) + aoff and realigns them by sft.
    uint64_t aoff[X == 2 ? G * OS : 1];
    uint32_t sft[X == 2 ? G * OS : 1];
    uint32_t mask[MW][NS];     // [w][slot]: outlier bits of frames g*F + 32w + b
    float lmax[NS];            // [slot]: the largest outlier distance (valid when lmaxi != ~0)
    uint32_t lmaxi[NS];        // [slot]: its first frame
    float med[4][P];
    float iqr_inv[4][P];
    uint32_t s0[4][P];       // frame 0 samples
    uint32_t sum_all[4][P];  // average modes
    uint32_t sum_out[4][P];
    __device__ static __forceinline__ uint32_t oidx(uint32_t t) { return t + 2u * (t / (uint32_t)F); }
    __device__ static __forceinline__ uint32_t rot(uint32_t g) {
        return X == 1 ? 4u * g : (g & 1u) + (g >> 1) * (64u / (uint32_t)G);
    }
    __device__ static __forceinline__ uint32_t slot(uint32_t g, uint32_t p) {
        return g * (uint32_t)P + ((p + rot(g)) & (uint32_t)(P - 1));
    }
};

// Fills the frame-offset table of a tile (all 256 lanes take part): byte offsets of the
// frame positions, the last one repeated up to `count` positions, at oidx(t) of a table
// with blocks of F entries.
template <int B, int F>
__device__ __forceinline__ void load_offsets(uint64_t* off, uint32_t count, const uint64_t* __restrict__ pos_off,
                                             uint32_t n) {
    for (uint32_t i = threadIdx.x; i < count; i += 256) off[i + 2u * (i / (uint32_t)F)] = pos_off[min(i, n - 1)] * B;
}

// load_offsets plus the realignment tables of a two-pixel tile (TileShared::aoff, sft); r is
// the tile row's first sample address modulo 4.
template <int B, int F>
__device__ __forceinline__ void load_offsets_pair(uint64_t* off, uint64_t* aoff, uint32_t* sft, uint32_t count,
                                                  const uint64_t* __restrict__ pos_off, uint32_t n, uint32_t r) {
    for (uint32_t i = threadIdx.x; i < count; i += 256) {
        const uint32_t j = i + 2u * (i / (uint32_t)F);
        const uint64_t o = pos_off[min(i, n - 1)] * B;
        off[j] = o;
        aoff[j] = (o + r) & ~(uint64_t)3;
        sft[j] = (uint32_t)(o + r) & 3u;
    }
}

// Loads the F frames of one lane block into raw[e*C + c][j/2] (pixel e of the lane, frame j
// in half j&1); base is the lane's first pixel in stack slot 0 and off the block's F frame
// byte offsets.
// Two pixels per lane (16-bit stacks) take one 16-byte load per frame from the dword
// boundary at or below the pair's first sample (3 channels: the pair's 12 bytes start 2-byte
// aligned, so the load reads up to 4 bytes past them and v_alignbyte realigns by the address'
// low bits).  All F loads are issued before the first is used, so a lane keeps its whole
// block in flight; the byte shifts are read from the tile's shift table afterwards rather
// than held in registers meanwhile.
template <int B, int C, int F, int X>
__device__ __forceinline__ void load_block(const uint8_t* __restrict__ base, const uint64_t* off,
                                           uint32_t (&raw)[X * C][F / 2], const uint64_t* aoff = nullptr,
                                           const uint32_t* sft = nullptr) {
    static_assert(X == 1 || (X == 2 && B == 2 && C >= 3), "two pixels per lane: 16-bit, 3 or 4 channels");
    if constexpr (X == 2) {
        typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
        typedef const __attribute__((address_space(1))) u32x4 gu4;
        typedef const volatile __attribute__((address_space(3))) uint32_t lu32;
        uint32_t q[F][4];
        const uintptr_t base_al = (uintptr_t)base & ~(uintptr_t)3;
#pragma unroll
        for (int t = 0; t < F; ++t) {
            const u32x4 v = *(gu4*)(base_al + aoff[t]);
            q[t][0] = v.x;
            q[t][1] = v.y;
            q[t][2] = v.z;
            q[t][3] = v.w;
        }
        __builtin_amdgcn_sched_barrier(0);
        lu32* sft32 = (lu32*)sft;
#pragma unroll
        for (int i = 0; i < F / 2; ++i) {
            uint32_t a[4], b[4];
            if constexpr (C == 3) {
                const uint32_t s0 = sft32[2 * i], s1 = sft32[2 * i + 1];
#pragma unroll
                for (int h = 0; h < 3; ++h) {  // [c0|c1], [c2|c0'], [c1'|c2'] of pixels 0 and 1
                    a[h] = __builtin_amdgcn_alignbyte(q[2 * i][h + 1], q[2 * i][h], s0);
                    b[h] = __builtin_amdgcn_alignbyte(q[2 * i + 1][h + 1], q[2 * i + 1][h], s1);
                }
            } else {
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    a[h] = q[2 * i][h];
                    b[h] = q[2 * i + 1][h];
                }
            }
#pragma unroll
            for (int h = 0; h < C; ++h) {
                raw[2 * h][i] = __builtin_amdgcn_perm(b[h], a[h], 0x05040100u);
                raw[2 * h + 1][i] = __builtin_amdgcn_perm(b[h], a[h], 0x07060302u);
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < F / 2; ++i) {
        const uint8_t* p0 = base + off[2 * i];
        const uint8_t* p1 = base + off[2 * i + 1];
        if constexpr (B == 2 && C == 3) {
            uint32_t d0, d1;  // channels 0-1 of each frame (2-byte aligned dword)
            __builtin_memcpy(&d0, __builtin_assume_aligned(p0, 2), 4);
            __builtin_memcpy(&d1, __builtin_assume_aligned(p1, 2), 4);
            raw[0][i] = __builtin_amdgcn_perm(d1, d0, 0x05040100u);
            raw[1][i] = __builtin_amdgcn_perm(d1, d0, 0x07060302u);
            uint32_t e0, e1;  // channels 1-2 of each frame
            __builtin_memcpy(&e0, __builtin_assume_aligned(p0 + 2, 2), 4);
            __builtin_memcpy(&e1, __builtin_assume_aligned(p1 + 2, 2), 4);
            raw[2][i] = __builtin_amdgcn_perm(e1, e0, 0x07060302u);
        } else if constexpr (B == 2 && C == 4) {
            const uint2 q0 = *reinterpret_cast<const uint2*>(p0);
            const uint2 q1 = *reinterpret_cast<const uint2*>(p1);
            raw[0][i] = __builtin_amdgcn_perm(q1.x, q0.x, 0x05040100u);
            raw[1][i] = __builtin_amdgcn_perm(q1.x, q0.x, 0x07060302u);
            raw[2][i] = __builtin_amdgcn_perm(q1.y, q0.y, 0x05040100u);
            raw[3][i] = __builtin_amdgcn_perm(q1.y, q0.y, 0x07060302u);
        } else if constexpr (B == 1 && C == 4) {
            const uint32_t d0 = *reinterpret_cast<const uint32_t*>(p0);
            const uint32_t d1 = *reinterpret_cast<const uint32_t*>(p1);
#pragma unroll
            for (int c = 0; c < 4; ++c)  // byte c of each frame into the low byte of each half
                raw[c][i] = __builtin_amdgcn_perm(d1, d0, 0x0C000C00u | ((uint32_t)(4 + c) << 16) | (uint32_t)c);
        } else if constexpr (B == 1 && C == 3) {
            uint16_t w0, w1;  // channels 0-1 of each frame
            __builtin_memcpy(&w0, p0, 2);
            __builtin_memcpy(&w1, p1, 2);
            raw[0][i] = __builtin_amdgcn_perm((uint32_t)w1, (uint32_t)w0, 0x0C040C00u);
            raw[1][i] = __builtin_amdgcn_perm((uint32_t)w1, (uint32_t)w0, 0x0C050C01u);
            us2 v;
            v.x = p0[2];
            v.y = p1[2];
            raw[2][i] = as_u32(v);
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                us2 v;
                v.x = (unsigned short)load_sample<B>(p0, c);
                v.y = (unsigned short)load_sample<B>(p1, c);
                raw[c][i] = as_u32(v);
            }
        }
    }
}

// fn(std::integral_constant<int, I>) for I = 0 .. N-1.
template <int I, int N>
struct StaticFor {
    template <typename Fn>
    __device__ static __forceinline__ void run(Fn&& fn) {
        if constexpr (I < N) {
            fn(std::integral_constant<int, I>{});
            StaticFor<I + 1, N>::run(fn);
        }
    }
};

// Slots per pixel of the outlier distances a one-lane-per-pixel tile keeps for the average
// mode's finish (odist below).
constexpr int kOdistSlots = 16;

// Finishes pixel p of a tile from the lanes' summaries in LDS: background and outlier
// selection / blending (src/chrono.rs:290-493), output and statistics.  odist (nullable):
// the classification's squared distances of pixel p's outliers, slot s * P + p holding the
// s-th outlier counted from the last frame, for up to kOdistSlots outliers; with it the
// average mode takes the distances from there instead of re-reading the samples.
template <int B, int C, int G, int F, int X, bool ABS>
__device__ __forceinline__ void finish_pixel(const OutlierK& k, uint32_t bx, uint32_t y, uint32_t p,
                                             const TileShared<C, G, F, X>& sh, const float* odist = nullptr) {
    using TS = TileShared<C, G, F, X>;
    constexpr int P = TS::P;
    constexpr int MW = TS::MW;
    constexpr uint32_t MAXV = B == 1 ? 255u : 65535u;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    // outliers whose samples the finish re-reads together: eight for the two-pixel tiles
    // (same registers as four, 0.4% faster on the 4K stack), one for the four-pixel-per-lane
    // tiles (a batch pushes their kernels into private memory)
    constexpr int KB = X >= 4 ? 1 : (X == 2 ? 8 : 4);
    const uint32_t N = k.n;
    const uint32_t x = bx * P + p;
    if (x >= k.out_w) return;
    const uint64_t pix = ((uint64_t)y * k.stack_w + x) * C;
    float med[4], iqr_inv[4], wi[4], sg[4];
    const bool want_iqr = !ABS || k.stats != nullptr;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        med[c] = sh.med[c][p];
        iqr_inv[c] = want_iqr ? sh.iqr_inv[c][p] : 0.0f;
        wi[c] = ABS ? k.w[c] : k.w[c] * iqr_inv[c];
        sg[c] = signum_f(k.w[c]);
    }
    uint32_t gm[G][MW];
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int w = 0; w < MW; ++w) gm[q][w] = sh.mask[w][TS::slot(q, p)];

    uint32_t num = 0, first_idx = NONE, last_idx = 0, first_non = NONE;
#pragma unroll
    for (int q = 0; q < G; ++q) {
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            const uint32_t mm = gm[q][w];
            const uint32_t base = (uint32_t)q * F + 32u * w;
            num += __builtin_popcount(mm);
            if (first_idx == NONE && mm) first_idx = base + __builtin_ctz(mm);
            if (mm) last_idx = base + 31 - __builtin_clz(mm);
            // valid non-outlier frames of this word: a word holds min(32, F - 32w) frames
            const int32_t span = F - 32 * w < 32 ? F - 32 * w : 32;
            const int32_t nw = min((int32_t)N - (int32_t)base, span);
            const uint32_t vm = nw >= 32 ? 0xFFFFFFFFu : (nw <= 0 ? 0u : (1u << nw) - 1u);
            const uint32_t non = vm & ~mm;
            if (first_non == NONE && non) first_non = base + __builtin_ctz(non);
        }
    }
    float max_d = 0.0f;
    uint32_t max_index = 0;
    if (k.outlier == 2 || k.stats != nullptr) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const uint32_t li = sh.lmaxi[TS::slot(q, p)];
            const float lm = sh.lmax[TS::slot(q, p)];
            if (li != NONE && lm > max_d) { max_d = lm; max_index = li; }
        }
    }
    if (first_idx == NONE) first_idx = 0;

    constexpr uint32_t maxv = MAXV;
    auto sample = [&](uint32_t t, uint32_t (&s)[4]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) s[ch] = load_sample<B>(k.stack + sh.off[TS::oidx(t)], pix + (uint64_t)ch);
    };
    auto dist_of = [&](uint32_t t, uint32_t (&s)[4]) -> float {
        sample(t, s);
        return frame_dist<C>(s, med, wi, sg);
    };
    const uint64_t pix_id = (uint64_t)y * k.out_w + x;
    uint32_t pixel[4] = {0, 0, 0, 0};
    bool warning = false;
    uint32_t bg_index = 0;
    uint32_t s[4] = {0, 0, 0, 0};
    if (k.background == 2) {
        float mean[4];
#pragma unroll
        for (int c = 0; c < C; ++c) mean[c] = (float)sh.sum_all[c][p] / (float)N;
        if (num == 1) {
            sample(first_idx, s);
#pragma unroll
            for (int c = 0; c < C; ++c)
                pixel[c] = sat_cast(roundf(mean[c] * ((float)N / (float)(N - 1)) - (float)s[c] / (float)N), maxv);
        } else if (num > 1) {
#pragma unroll
            for (int c = 0; c < C; ++c)
                pixel[c] = sat_cast(roundf(mean[c] * ((float)N / (float)(N - num)) - (float)sh.sum_out[c][p] / (float)N), maxv);
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) pixel[c] = sat_cast(roundf(mean[c]), maxv);
        }
    } else if (k.background == 3) {
#pragma unroll
        for (int c = 0; c < C; ++c) pixel[c] = sat_cast(roundf(med[c]), maxv);
    } else {
        if (k.background == 0) {
            if (num == N) { bg_index = 0; warning = true; }
            else bg_index = num == 0 ? 0 : first_non;
        } else {
            if (num == 0 || num == N) {
                bg_index = rng_range(k.seed, k.pix_base + pix_id, 0, N);
                warning = num == N;
            } else {
                uint32_t pos = rng_range(k.seed, k.pix_base + pix_id, 0, N - num);
                uint32_t j = num;
#pragma unroll
                for (int q = G - 1; q >= 0; --q) {
#pragma unroll
                    for (int w = MW - 1; w >= 0; --w) {
                        uint32_t mm = gm[q][w];
                        while (mm) {
                            const uint32_t b = 31 - __builtin_clz(mm);
                            mm &= ~(1u << b);
                            const uint32_t a = (uint32_t)q * F + (uint32_t)w * 32 + b;
                            --j;
                            const uint32_t bb = N - 1 - j;
                            if (pos == a) pos = bb;
                            else if (pos == bb) pos = a;
                        }
                    }
                }
                bg_index = pos;
            }
        }
        if (bg_index == 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) s[c] = sh.s0[c][p];
        } else {
            sample(bg_index, s);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) pixel[c] = s[c];
    }

    uint32_t blend_byte = 0;
    if (num > 0) {
        if (num == 1) {
            const float d = dist_of(first_idx, s);
            const float fade = fade_pick(k.fade, (int32_t)first_idx, (int32_t)N, k.frame_offset);
            const float blend = fade * blend_value(k.th, sqrtf(d));
            blend_into_px<C>(pixel, s, blend, maxv);
            blend_byte = sat_cast(roundf(blend * 255.0f), 255u);
        } else if (k.outlier == 4 || k.outlier == 5) {
            float pix_new[4] = {0, 0, 0, 0};
            float blend_inv = 1.0f;
#pragma unroll
            for (int c = 0; c < C; ++c) pix_new[c] = (float)pixel[c];
            // outliers in frame order (forward) or reverse order (backward); the direction is a
            // template constant so that every mask word is a fixed register
            auto walk = [&](auto fwd_t) {
                constexpr bool FWD = decltype(fwd_t)::value;
#pragma unroll
                for (int qq = 0; qq < G; ++qq) {
                    const int q = FWD ? qq : G - 1 - qq;
#pragma unroll
                    for (int ww = 0; ww < MW; ++ww) {
                        const int w = FWD ? ww : MW - 1 - ww;
                        uint32_t mm = gm[q][w];
                        while (mm) {
                            // up to KB outliers at a time: all their samples are requested
                            // before the first is blended, so the re-reads' latencies overlap;
                            // the blends still run one after another in frame order
                            uint32_t tb[KB], sb[KB][4];
                            bool vb[KB];
#pragma unroll
                            for (int u = 0; u < KB; ++u) {
                                vb[u] = mm != 0u;
                                const uint32_t b =
                                    vb[u] ? (FWD ? (uint32_t)__builtin_ctz(mm) : 31 - __builtin_clz(mm)) : 0u;
                                if (vb[u]) mm &= ~(1u << b);
                                tb[u] = vb[u] ? (uint32_t)q * F + (uint32_t)w * 32 + b : tb[0];
                            }
#pragma unroll
                            for (int u = 0; u < KB; ++u) sample(tb[u], sb[u]);
#pragma unroll
                            for (int u = 0; u < KB; ++u) {
                                if (!vb[u]) break;
                                const uint32_t t = tb[u];
                                const float d = frame_dist<C>(sb[u], med, wi, sg);
                                const float fade = fade_pick(k.fade, (int32_t)t, (int32_t)N, k.frame_offset);
                                const float blend = fade * blend_value(k.th, sqrtf(d));
                                blend_into_f32<C>(pix_new, sb[u], blend);
                                blend_inv *= 1.0f - blend;
                            }
                        }
                    }
                }
            };
            if (k.outlier == 4) walk(std::true_type{});
            else walk(std::false_type{});
#pragma unroll
            for (int c = 0; c < C; ++c) pixel[c] = sat_cast(roundf(pix_new[c]), maxv);
            blend_byte = sat_cast(roundf((1.0f - blend_inv) * 255.0f), 255u);
        } else {
            uint32_t sel = 0;
            float dist;
            if (k.outlier == 3) {
                float mean_dist = 0.0f;
                if (odist != nullptr && num <= (uint32_t)kOdistSlots) {
                    // frame order: the outlier of the lowest frame sits in the last slot
                    for (uint32_t i = num; i-- > 0;) mean_dist += sqrtf(odist[i * P + p]);
                } else {
#pragma unroll
                    for (int q = 0; q < G; ++q) {
#pragma unroll
                        for (int w = 0; w < MW; ++w) {
                            uint32_t mm = gm[q][w];
                            while (mm) {  // KB re-reads in flight at a time, summed in frame order
                                uint32_t tb[KB], sb[KB][4];
                                bool vb[KB];
#pragma unroll
                                for (int u = 0; u < KB; ++u) {
                                    vb[u] = mm != 0u;
                                    const uint32_t b = vb[u] ? (uint32_t)__builtin_ctz(mm) : 0u;
                                    mm &= vb[u] ? mm - 1 : mm;
                                    tb[u] = vb[u] ? (uint32_t)q * F + (uint32_t)w * 32 + b : tb[0];
                                }
#pragma unroll
                                for (int u = 0; u < KB; ++u) sample(tb[u], sb[u]);
#pragma unroll
                                for (int u = 0; u < KB; ++u) {
                                    if (!vb[u]) break;
                                    mean_dist += sqrtf(frame_dist<C>(sb[u], med, wi, sg));
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < C; ++c) s[c] = sat_cast(roundf((float)sh.sum_out[c][p] / (float)num), maxv);
                dist = mean_dist / (float)num;
            } else {
                sel = k.outlier == 0 ? first_idx : (k.outlier == 1 ? last_idx : max_index);
                dist = sqrtf(dist_of(sel, s));
            }
            const float fade = fade_pick(k.fade, (int32_t)sel, (int32_t)N, k.frame_offset);
            const float blend = fade * blend_value(k.th, dist);
            blend_into_px<C>(pixel, s, blend, maxv);
            blend_byte = sat_cast(roundf(blend * 255.0f), 255u);
        }
    }

    const uint64_t o = ((uint64_t)y * k.out_w + x) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        store_sample<B>(k.out_px, o + c, pixel[c]);
        if (k.out_mask) k.out_mask[o + c] = c < 3 ? (uint8_t)blend_byte : (uint8_t)255;
    }
    if (warning && k.warnings) atomicAdd(k.warnings, 1ull);
    if (k.stats) {
        PixelStats st;
#pragma unroll
        for (int i = 0; i < 4; ++i) { st.median[i] = med[i]; st.iqr_inv[i] = iqr_inv[i]; }
        st.num_outliers = num;
        st.max_index = max_index;
        st.max_dist_sq = max_d;
        st.bg_index = bg_index;
        st.warning = warning ? 1u : 0u;
        st.pad = 0;
        k.stats[pix_id] = st;
    }
}

// One tile: P = X*256/G pixels of row y starting at pixel bx * P; lane (pl, g) holds frame
// block g of pixels bx*P + pl*X + e, e < X.  sh.off must hold the tile's frame offsets
// (load_offsets) and be visible to all lanes.  Every lane of the workgroup must call this
// (it synchronises once).  With X = 2 the lanes read up to 16 bytes past their pixels'
// samples (load12_a2, and pixel out_w of a row when out_w is odd); the launcher checks that
// those bytes lie inside the stack allocation.  lane: threadIdx.x.
template <int B, int C, int G, int F, int X, bool ABS>
__device__ __forceinline__ void outlier_tile(const OutlierK& k, uint32_t bx, uint32_t y,
                                             TileShared<C, G, F, X>& sh, uint32_t lane) {
    static_assert(F % 4 == 0 && F <= 64 && G <= 8 && (X == 1 || X == 2), "tile shape");
    using TS = TileShared<C, G, F, X>;
    constexpr int P = TS::P;
    constexpr int MW = TS::MW;
    constexpr int NW = F / 2;
    constexpr uint32_t MAXV = B == 1 ? 255u : 65535u;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    const uint32_t N = k.n;

    const uint32_t g = lane % G;
    const uint32_t t0 = g * F;
    const int32_t ntot = G * F;
    {
        const uint32_t pl = lane / G;
        const uint32_t x0 = bx * P + pl * X;
        const uint64_t pix = ((uint64_t)y * k.stack_w + (x0 < k.out_w ? x0 : 0u)) * C;

        // ---- the lane's frame block; pads above N are set to MAXV
        uint32_t raw[X * C][NW];
        if constexpr (X == 2) load_block<B, C, F, X>(k.stack + pix * B, sh.off, raw, sh.aoff + g * TS::OS, sh.sft + g * TS::OS);
        else load_block<B, C, F, X>(k.stack + pix * B, sh.off + g * TS::OS, raw);
        if (N < (uint32_t)(G * F)) {  // grid-uniform: only stacks that do not fill the tile
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                const uint32_t t = t0 + 2 * i;
#pragma unroll
                for (int c = 0; c < X * C; ++c) {
                    uint32_t r = raw[c][i];
                    r = t + 1 >= N ? (r & 0xFFFFu) | (MAXV << 16) : r;
                    r = t >= N ? MAXV | (MAXV << 16) : r;
                    raw[c][i] = r;
                }
            }
        }

        // ---- medians / inverse IQR per pixel and channel
        const RankPlan plan = rank_plan_core(N, ABS);
        const uint32_t real = N > t0 ? min(N - t0, (uint32_t)F) : 0u;  // real frames in this block
        float med[X][4], iqr_inv[X][4];
        StaticFor<0, X>::run([&](auto e_t) {
            constexpr int E = decltype(e_t)::value;
            const uint32_t (&re)[C][NW] = *reinterpret_cast<const uint32_t(*)[C][NW]>(&raw[E * C][0]);
            uint32_t m0[C], m1[C], a0[C], a1[C], b0[C], b1[C];
            if constexpr (!ABS && B == 1) {
                rank_triples_u8<G, NW, C>(re, plan, ntot, real, m0, m1, a0, a1, b0, b1);
            } else {
                rank_pairs<B, G, NW, C, ABS>(re, plan.m0, plan.m1, ntot, real, m0, m1);
                if constexpr (!ABS) {
                    rank_pairs<B, G, NW, C, false>(re, plan.a0, plan.a1, ntot, real, a0, a1);
                    rank_pairs<B, G, NW, C, false>(re, plan.b0, plan.b1, ntot, real, b0, b1);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                med[E][c] = 0.0f;
                iqr_inv[E][c] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float mc, ic;
                if constexpr (ABS) stats_from_values(plan, m0[c], m1[c], 0, 0, 0, 0, true, &mc, &ic);
                else stats_from_values(plan, m0[c], m1[c], a0[c], a1[c], b0[c], b1[c], false, &mc, &ic);
                // channels without weight keep zero statistics
                med[E][c] = k.w[c] == 0.0f ? 0.0f : mc;
                iqr_inv[E][c] = k.w[c] == 0.0f ? 0.0f : ic;
            }
        });

        // ---- classification over the lane's frames
        float wi[X][4], sg[4];
        bool unit = ABS, pos = true;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sg[c] = signum_f(k.w[c]);
#pragma unroll
            for (int e = 0; e < X; ++e) wi[e][c] = ABS ? k.w[c] : k.w[c] * iqr_inv[e][c];
            if (c < C) {
                unit = unit && k.w[c] == 1.0f;
                pos = pos && !(sg[c] < 0.0f);
            }
        }
        const float tsq = k.th.min * k.th.min;
        const bool want_max = k.outlier == 2 || k.stats != nullptr;
        const bool want_sums = k.background == 2 || k.outlier == 3;
        const bool full_tile = N == (uint32_t)(G * F);
        uint32_t m[X][MW];
        float lmax[X];
        uint32_t lmaxi[X];
        uint32_t sum_all[X][4], sum_out[X][4];
#pragma unroll
        for (int e = 0; e < X; ++e) {
#pragma unroll
            for (int w = 0; w < MW; ++w) m[e][w] = 0;
            lmax[e] = tsq;
            lmaxi[e] = NONE;
#pragma unroll
            for (int c = 0; c < 4; ++c) sum_all[e][c] = sum_out[e][c] = 0;
        }

        // Frames run from last to first so that the mask shifts in bit j for frame j; the
        // largest distance keeps the lowest frame among ties (>=).  lmax starts at the
        // threshold, so `d >= lmax` also implies that the frame is an outlier.
        auto classify = [&](auto e_t, auto unit_t, auto pos_t, auto max_t, auto sums_t, auto check_t) {
            constexpr int E = decltype(e_t)::value;
            constexpr bool UNIT = decltype(unit_t)::value, POS = decltype(pos_t)::value;
            constexpr bool MAXT = decltype(max_t)::value, SUMS = decltype(sums_t)::value;
            constexpr bool CHECK = decltype(check_t)::value;
#pragma unroll
            for (int j = F - 1; j >= 0; --j) {
                uint32_t xs[4];
                float d = 0.0f;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const uint32_t w32 = raw[E * C + c][j >> 1];
                    xs[c] = (j & 1) ? (w32 >> 16) : (w32 & 0xFFFFu);
                    const float diff = med[E][c] - (float)xs[c];
                    const float a = UNIT ? diff : wi[E][c] * diff;
                    float term = a * a;
                    if constexpr (!POS) term = sg[c] * term;
                    d = c == 0 ? term : d + term;  // 0 + term0 == term0 up to the sign of zero
                }
                const bool o = d >= tsq;  // pads are masked out after the loop
                m[E][j >> 5] = (m[E][j >> 5] << 1) | (o ? 1u : 0u);
                bool in = true;
                if constexpr (CHECK) in = t0 + (uint32_t)j < N;
                if constexpr (MAXT) {
                    if (in && d >= lmax[E]) { lmax[E] = d; lmaxi[E] = t0 + (uint32_t)j; }
                }
                if constexpr (SUMS && CHECK) {
#pragma unroll
                    for (int c = 0; c < C; ++c) sum_out[E][c] += (o && in) ? xs[c] : 0u;
                }
            }
            if constexpr (SUMS && !CHECK) {
                // every frame is real: the outliers' sums from the mask, two frames per
                // v_dot2_u32_u16 (the mask's two bits spread to one 0/1 half per frame)
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    const uint32_t b = (m[E][i >> 4] >> (2 * (i & 15))) & 3u;
                    const uint32_t sel = (b & 1u) | ((b & 2u) << 15);
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        sum_out[E][c] = __builtin_amdgcn_udot2(as_us2(raw[E * C + c][i]), as_us2(sel), sum_out[E][c], false);
                }
            }
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        StaticFor<0, X>::run([&](auto e_t) {
            constexpr int E = decltype(e_t)::value;
            // Bound check: with non-negative weights every frame's f32 distance is at most
            // the same f32 expression on the per-channel maximum deviation from the median
            // (correct rounding is monotone), so when that bound is below the threshold no
            // frame of this lane can be an outlier and the per-frame pass is skipped.
            bool may_outlier = true;
            if (pos && full_tile) {
                float bound = 0.0f;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    us2 mn = {0xFFFF, 0xFFFF}, mx = {0, 0};
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        mn = __builtin_elementwise_min(mn, as_us2(raw[E * C + c][i]));
                        mx = __builtin_elementwise_max(mx, as_us2(raw[E * C + c][i]));
                    }
                    const float lo = (float)min((uint32_t)mn.x, (uint32_t)mn.y);
                    const float hi = (float)max((uint32_t)mx.x, (uint32_t)mx.y);
                    const float dev = fmaxf(med[E][c] - lo, hi - med[E][c]);
                    const float a = unit ? dev : wi[E][c] * dev;
                    bound = c == 0 ? a * a : bound + a * a;
                }
                may_outlier = bound >= tsq;
            }
            if (may_outlier) {
                if (!full_tile) classify(e_t, F_{}, F_{}, T_{}, T_{}, T_{});
                else if (want_max) {
                    if (want_sums) classify(e_t, F_{}, F_{}, T_{}, T_{}, F_{});
                    else if (unit) classify(e_t, T_{}, T_{}, T_{}, F_{}, F_{});
                    else if (pos) classify(e_t, F_{}, T_{}, T_{}, F_{}, F_{});
                    else classify(e_t, F_{}, F_{}, T_{}, F_{}, F_{});
                } else {
                    if (want_sums) {
                        if (pos) classify(e_t, F_{}, T_{}, F_{}, T_{}, F_{});
                        else classify(e_t, F_{}, F_{}, F_{}, T_{}, F_{});
                    } else if (unit) classify(e_t, T_{}, T_{}, F_{}, F_{}, F_{});
                    else if (pos) classify(e_t, F_{}, T_{}, F_{}, F_{}, F_{});
                    else classify(e_t, F_{}, F_{}, F_{}, F_{}, F_{});
                }
            }
        });
        if (want_sums) {
            // all-frame sums through v_sad_u16 against zero; pads (MAXV) are taken back out
            const uint32_t npad = t0 + F > N ? min(t0 + F - N, (uint32_t)F) : 0u;
#pragma unroll
            for (int e = 0; e < X; ++e)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int i = 0; i < NW; ++i) acc = __builtin_amdgcn_sad_u16(raw[e * C + c][i], 0u, acc);
                    sum_all[e][c] = (uint32_t)grp_sum<G>((int32_t)(acc - npad * MAXV));
                    sum_out[e][c] = (uint32_t)grp_sum<G>((int32_t)sum_out[e][c]);
                }
        }

        // ---- leave the lane's summaries and the pixels' statistics in LDS
        {
            const int32_t nv = min((int32_t)N - (int32_t)t0, (int32_t)F);  // valid frames in this block
#pragma unroll
            for (int e = 0; e < X; ++e) {
                const uint32_t sl = TS::slot(g, pl * X + e);
#pragma unroll
                for (int w = 0; w < MW; ++w) {
                    const int32_t nw = nv - 32 * w;
                    const uint32_t vm = nw >= 32 ? 0xFFFFFFFFu : (nw <= 0 ? 0u : (1u << nw) - 1u);
                    sh.mask[w][sl] = m[e][w] & vm;
                }
                if (want_max) {
                    sh.lmax[sl] = lmax[e];
                    sh.lmaxi[sl] = lmaxi[e];
                }
            }
        }
        if (g == 0) {
            // only what the finish reads for these options
            const bool want_iqr = !ABS || k.stats != nullptr;
            const bool want_s0 = k.background < 2;
#pragma unroll
            for (int e = 0; e < X; ++e) {
                const uint32_t p = pl * X + e;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    sh.med[c][p] = med[e][c];
                    if (want_iqr) sh.iqr_inv[c][p] = iqr_inv[e][c];
                }
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    if (want_s0) sh.s0[c][p] = raw[e * C + c][0] & 0xFFFFu;
                    if (want_sums) {
                        sh.sum_all[c][p] = sum_all[e][c];
                        sh.sum_out[c][p] = sum_out[e][c];
                    }
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t p = lane; p < (uint32_t)P; p += 256) finish_pixel<B, C, G, F, X, ABS>(k, bx, y, p, sh);
}

// Workgroup b of nb goes to XCD b % 8 in dispatch order; renumbering the tiles gives each
// XCD one contiguous run, so neighbouring tiles (which share cache lines at their edges)
// meet in the same L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / 8u, r = nb % 8u, xcd = b % 8u, idx = b / 8u;
    return xcd * q + min(xcd, r) + idx;
}

// Waves per SIMD the register allocator is asked to fit without spilling: 3-channel 32-frame
// tiles fit five (absolute thresholds) or four (relative thresholds: three rank searches),
// 4-channel tiles one fewer; two pixels per lane double the frame registers (three waves,
// two for 4 channels or relative thresholds); the 64-frame tiles get the whole register file.
constexpr int waves_hint(int F, int C, bool absolute, int X = 1) {
    return F > 32 ? 1 : X == 2 ? (absolute && C == 3 ? 3 : 2) : (absolute ? 5 : 4) - (C > 3 ? 1 : 0);
}

template <int B, int C, int G, int F, int X, bool ABS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint(F, C, ABS, X))))
outlier_fast_kernel(OutlierK k, uint32_t tiles_x) {
    __shared__ TileShared<C, G, F, X> sh;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t y = tile / tiles_x;
    if constexpr (X == 2) {
        const uint32_t r = (uint32_t)((uintptr_t)k.stack + (uint64_t)y * k.stack_w * C * B) & 3u;
        load_offsets_pair<B, F>(sh.off, sh.aoff, sh.sft, G * F, k.pos_off, k.n, r);
    } else {
        load_offsets<B, F>(sh.off, G * F, k.pos_off, k.n);
    }
    __syncthreads();
    outlier_tile<B, C, G, F, X, ABS>(k, tile % tiles_x, y, sh, threadIdx.x);
}

}  // namespace cp
