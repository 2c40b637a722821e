// k_outlier_fast.h -- register-resident outlier composite (the hot path).
//
// Layout.  A workgroup of 256 lanes covers P = 256/G consecutive output pixels of one row.
// The G lanes of a pixel (adjacent lanes, one DPP quad or half-row) each hold a contiguous
// block of F frames x C channels in VGPRs, two samples per VGPR (16-bit halves, frames
// 2j and 2j+1), so the stack tile is read from HBM exactly once and stays on chip for the
// order-statistic pass, the classification pass and the blend.  Frame t's element offset
// (stack slot + crop origin) comes from a per-workgroup LDS table.
//
// Order statistics (src/chrono.rs:228-255 sorts each weighted channel; any exact selection
// gives the same values).  The r-th smallest value is found MSB-first: keep the largest
// prefix v with #{y < v} <= r.  Counts come from the sum of absolute differences:
//     sum_y |y - v| - sum_y |y - (v-1)| = 2 #{y < v} - n
// so one search step costs two v_sad_u8 per four byte-samples plus a G-lane DPP reduction.
// 8-bit stacks need one 8-step level; 16-bit stacks run one level on the high bytes and
// one on the low bytes clamped into the winning high-byte bucket (values below map to 0,
// above to 255, which leaves every count #{y < v}, v in [1,255], exact).  The neighbour
// order statistic of an even-sized median (and the quartile interpolation partners) is
// either the same value (when #{x <= v} covers it) or min{x > v} (packed u16 min).
// Frames beyond N are padded with the maximum sample value, which never changes the ranks
// below N.
//
// Classification (src/chrono.rs:257-288) runs over the lane's frames in order with the
// reference's f32 operation order; outliers become bits of a per-lane mask.  The group then
// reduces counts, the first maximum (extreme mode) and exact integer sums (average modes;
// the host routes stacks whose sums could exceed 2^24 to the generic kernel).  One lane per
// pixel finishes: background and outlier selection / blending (src/chrono.rs:290-493),
// re-reading from L2 only the few outlier samples an order-dependent blend needs.
#pragma once
#include <type_traits>

#include "device_common.h"
#include "k_outlier_generic.h"

namespace cp {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 as_us2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t as_u32(us2 v) { return __builtin_bit_cast(uint32_t, v); }

// Sum / min / max over the G adjacent lanes of a pixel group (G in {1,2,4,8}); the DPP
// moves fold into v_add/v_min/v_max_u32_dpp.
template <int G>
__device__ __forceinline__ int32_t grp_sum(int32_t v) {
    if constexpr (G >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);   // xor 1
    if constexpr (G >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);   // xor 2
    if constexpr (G >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);  // half-row mirror
    return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_min_u32(uint32_t v) {
    if constexpr (G >= 2) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0xB1, 0xF, 0xF, false));
    if constexpr (G >= 4) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x4E, 0xF, 0xF, false));
    if constexpr (G >= 8) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x141, 0xF, 0xF, false));
    return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_max_u32(uint32_t v) {
    if constexpr (G >= 2) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    if constexpr (G >= 4) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    if constexpr (G >= 8) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    return v;
}

// #{bytes < v} over the group, v in [1, 255].
template <int G, int ND>
__device__ __forceinline__ uint32_t byte_count_lt(const uint32_t (&y)[ND], uint32_t v, int32_t ntot) {
    const uint32_t V = v * 0x01010101u, Vm = V - 0x01010101u;
    uint32_t a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        a = __builtin_amdgcn_sad_u8(y[i], V, a);
        b = __builtin_amdgcn_sad_u8(y[i], Vm, b);
    }
    return (uint32_t)(ntot + grp_sum<G>((int32_t)(a - b))) >> 1;
}

// Largest v in [0, 255] with #{bytes < v} <= k over the group's ND dwords per lane.  The
// candidate is kept replicated in all four bytes, so each step is two v_sad_u8 per dword,
// a G-lane reduction and a select.
template <int G, int ND>
__device__ __forceinline__ uint32_t byte_select(const uint32_t (&y)[ND], uint32_t k, int32_t ntot) {
    uint32_t lo = 0;
    const int32_t two_k = 2 * (int32_t)k;
#pragma unroll 1
    for (int bit = 7; bit >= 0; --bit) {
        const uint32_t V = lo | (0x01010101u << bit);
        const uint32_t Vm = V - 0x01010101u;
        uint32_t a = 0, b = 0;
#pragma unroll
        for (int i = 0; i < ND; ++i) {
            a = __builtin_amdgcn_sad_u8(y[i], V, a);
            b = __builtin_amdgcn_sad_u8(y[i], Vm, b);
        }
        // ntot + d = 2 #{y < v}
        if (ntot + grp_sum<G>((int32_t)(a - b)) <= two_k) lo = V;
    }
    return lo & 0xFFu;
}

// Exact r-th smallest sample of one 16-bit channel by two byte levels.
template <int G, int NW>
__device__ __forceinline__ uint32_t select_rank_u16(const uint32_t (&r)[NW], uint32_t k, int32_t ntot) {
    constexpr int ND = NW / 2;
    uint32_t y[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) y[i] = __builtin_amdgcn_perm(r[2 * i + 1], r[2 * i], 0x07050301u);
    const uint32_t H = byte_select<G, ND>(y, k, ntot);
    const us2 base = {(unsigned short)(H << 8), (unsigned short)(H << 8)};
    const us2 c255 = {255, 255};
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        us2 a = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i]), base), c255);
        us2 b = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i + 1]), base), c255);
        y[i] = __builtin_amdgcn_perm(as_u32(b), as_u32(a), 0x06040200u);
    }
    const uint32_t L = byte_select<G, ND>(y, k, ntot);
    return (H << 8) | L;
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

__device__ __forceinline__ us2 med3_us2(us2 a, us2 b, us2 c) {
    const us2 lo = __builtin_elementwise_min(a, b), hi = __builtin_elementwise_max(a, b);
    return __builtin_elementwise_max(lo, __builtin_elementwise_min(hi, c));
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

// Values at ranks (r0, r1), r1 in {r0, r0 + 1}, of one channel.
//  u8:  one byte level; the neighbour comes from one more byte count.
//  u16 (median): an 8-step byte search inside the window [lo, lo + 255] around the
//       estimate; the result is exact when it ends strictly inside (1..254), otherwise
//       the exact two-level search runs.  Group-uniform branches only.
template <int B, int G, int NW, bool WINDOW>
__device__ __forceinline__ void rank_pair(const uint32_t (&r)[NW], uint32_t r0, uint32_t r1, int32_t ntot,
                                          uint32_t& v0, uint32_t& v1) {
    constexpr int ND = NW / 2;
    constexpr uint32_t MAXV = B == 1 ? 255u : 65535u;
    uint32_t y[ND];
    uint32_t cle = 0;
    bool have_cle = false;
    if constexpr (B == 1) {
#pragma unroll
        for (int i = 0; i < ND; ++i) y[i] = __builtin_amdgcn_perm(r[2 * i + 1], r[2 * i], 0x06040200u);
        v0 = byte_select<G, ND>(y, r0, ntot);
        if (r1 != r0) {
            cle = v0 >= 255u ? (uint32_t)ntot : byte_count_lt<G, ND>(y, v0 + 1, ntot);
            have_cle = true;
        }
    } else if constexpr (WINDOW) {
        const uint32_t est = group_estimate<G, NW>(r);
        const uint32_t lo = min(est >= 127u ? est - 127u : 0u, 65535u - 255u);
        const us2 base = {(unsigned short)lo, (unsigned short)lo};
        const us2 c255 = {255, 255};
#pragma unroll
        for (int i = 0; i < ND; ++i) {
            us2 a = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i]), base), c255);
            us2 b = __builtin_elementwise_min(__builtin_elementwise_sub_sat(as_us2(r[2 * i + 1]), base), c255);
            y[i] = __builtin_amdgcn_perm(as_u32(b), as_u32(a), 0x06040200u);
        }
        const uint32_t ys = byte_select<G, ND>(y, r0, ntot);
        if (ys >= 1u && ys <= 254u) {
            v0 = lo + ys;
            if (r1 != r0) {
                cle = byte_count_lt<G, ND>(y, ys + 1, ntot);  // #{x < v0 + 1}
                have_cle = true;
            }
        } else {
            v0 = select_rank_u16<G, NW>(r, r0, ntot);
        }
    } else {
        v0 = select_rank_u16<G, NW>(r, r0, ntot);
    }
    if (r1 == r0) { v1 = v0; return; }
    if (!have_cle) cle = v0 >= MAXV ? (uint32_t)ntot : count_le<G, NW>(r, v0, ntot);
    v1 = cle >= r1 + 1 ? v0 : min_greater<G, NW>(r, v0);
}

template <int B, int C>
struct FastPixel {
    const OutlierK& k;
    const uint64_t* off;  // LDS offsets
    uint64_t pix;
    __device__ __forceinline__ void sample(uint32_t t, uint32_t (&s)[4]) const {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) s[ch] = load_sample<B>(k.stack, off[t] + pix + (uint64_t)ch);
    }
};

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

// One tile: P = 256/G pixels of row y starting at pixel bx * P.  s_off holds the element
// offsets of the N frame positions (LDS).  Lanes that do not finish a pixel return early.
template <int B, int C, int G, int F, bool ABS>
__device__ __forceinline__ void outlier_tile(const OutlierK& k, uint32_t bx, uint32_t y,
                                             const uint64_t* __restrict__ s_off) {
    static_assert(F % 4 == 0 && F <= 64 && G <= 8, "tile shape");
    constexpr int P = 256 / G;
    constexpr int NW = F / 2;
    constexpr uint32_t MAXV = B == 1 ? 255u : 65535u;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    const uint32_t N = k.n;

    const uint32_t g = threadIdx.x % G;
    const uint32_t p = threadIdx.x / G;
    const uint32_t x = bx * P + p;
    const bool px_ok = x < k.out_w;
    const uint64_t pix = ((uint64_t)y * k.stack_w + (px_ok ? x : 0u)) * C;
    const uint32_t t0 = g * F;
    const int32_t ntot = G * F;

    // ---- load the lane's frame block.  Loads are unconditional (frame index clamped to
    // N-1) so they issue back to back; pads above N are then set to MAXV.
    uint32_t raw[C][NW];
    {
        us2 rv[C][NW];
#pragma unroll
        for (int j = 0; j < F; ++j) {
            const uint32_t t = min(t0 + j, N - 1);
            const uint64_t base = s_off[t] + pix;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const unsigned short v = (unsigned short)load_sample<B>(k.stack, base + (uint64_t)c);
                if (j & 1) rv[c][j >> 1].y = v;
                else rv[c][j >> 1].x = v;
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int i = 0; i < NW; ++i) raw[c][i] = as_u32(rv[c][i]);
    }
    if (N < (uint32_t)(G * F)) {  // grid-uniform: only stacks that do not fill the tile
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const uint32_t t = t0 + 2 * i;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                uint32_t r = raw[c][i];
                r = t + 1 >= N ? (r & 0xFFFFu) | (MAXV << 16) : r;
                r = t >= N ? MAXV | (MAXV << 16) : r;
                raw[c][i] = r;
            }
        }
    }

    // ---- medians / inverse IQR per weighted channel
    const RankPlan plan = make_rank_plan(N, ABS);
    float med[4] = {0, 0, 0, 0}, iqr_inv[4] = {0, 0, 0, 0};
    // One code path serves every channel (a rolled loop keeps the kernel small); the
    // channel's registers are selected into `cur` with static register indices.
#pragma unroll 1
    for (int c = 0; c < C; ++c) {
        if (k.w[c] == 0.0f) continue;
        uint32_t cur[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) cur[i] = raw[0][i];
#pragma unroll
        for (int cc = 1; cc < C; ++cc)
#pragma unroll
            for (int i = 0; i < NW; ++i) cur[i] = c == cc ? raw[cc][i] : cur[i];
        uint32_t vals[6] = {0, 0, 0, 0, 0, 0};
        uint32_t m0, m1;
        rank_pair<B, G, NW, ABS>(cur, plan.m0, plan.m1, ntot, m0, m1);
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if constexpr (!ABS) {
            rank_pair<B, G, NW, false>(cur, plan.a0, plan.a1, ntot, a0, a1);
            rank_pair<B, G, NW, false>(cur, plan.b0, plan.b1, ntot, b0, b1);
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const uint32_t rk = plan.rank[i];
            uint32_t v = 0;
            if (rk == plan.m0) v = m0;
            else if (rk == plan.m1) v = m1;
            else if (rk == plan.a0) v = a0;
            else if (rk == plan.a1) v = a1;
            else if (rk == plan.b0) v = b0;
            else if (rk == plan.b1) v = b1;
            vals[i] = v;
        }
        float mc, ic;
        stats_from_ranks(plan, vals, ABS, &mc, &ic);
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            med[cc] = c == cc ? mc : med[cc];
            iqr_inv[cc] = c == cc ? ic : iqr_inv[cc];
        }
    }

    // ---- classification over the lane's frames
    float wi[4], sg[4];
    bool unit = ABS, pos = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        wi[c] = ABS ? k.w[c] : k.w[c] * iqr_inv[c];
        sg[c] = signum_f(k.w[c]);
        if (c < C) {
            unit = unit && k.w[c] == 1.0f;
            pos = pos && !(sg[c] < 0.0f);
        }
    }
    const float tsq = k.th.min * k.th.min;
    const bool want_max = k.outlier == 2 || k.stats != nullptr;
    const bool want_sums = k.background == 2 || k.outlier == 3;
    const bool full_tile = N == (uint32_t)(G * F);
    uint32_t m_lo = 0, m_hi = 0;
    float lmax = 0.0f;
    uint32_t lmaxi = NONE;
    uint32_t sum_all[4] = {0, 0, 0, 0}, sum_out[4] = {0, 0, 0, 0};

    // Bound check: with non-negative weights every frame's f32 distance is at most the
    // same f32 expression on the per-channel maximum deviation from the median (correct
    // rounding is monotone), so when that bound is below the threshold no frame of this
    // lane can be an outlier and the per-frame pass is skipped.
    bool may_outlier = true;
    if (pos && full_tile) {
        float bound = 0.0f;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            us2 mn = {0xFFFF, 0xFFFF}, mx = {0, 0};
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                mn = __builtin_elementwise_min(mn, as_us2(raw[c][i]));
                mx = __builtin_elementwise_max(mx, as_us2(raw[c][i]));
            }
            const float lo = (float)min((uint32_t)mn.x, (uint32_t)mn.y);
            const float hi = (float)max((uint32_t)mx.x, (uint32_t)mx.y);
            const float dev = fmaxf(med[c] - lo, hi - med[c]);
            const float a = unit ? dev : wi[c] * dev;
            bound = c == 0 ? a * a : bound + a * a;
        }
        may_outlier = bound >= tsq;
    }

    auto classify = [&](auto unit_t, auto pos_t, auto max_t, auto sums_t) {
        constexpr bool UNIT = decltype(unit_t)::value, POS = decltype(pos_t)::value;
        constexpr bool MAXT = decltype(max_t)::value, SUMS = decltype(sums_t)::value;
#pragma unroll
        for (int j = 0; j < F; ++j) {
            uint32_t xs[4];
            float d = 0.0f;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const uint32_t w32 = raw[c][j >> 1];
                xs[c] = (j & 1) ? (w32 >> 16) : (w32 & 0xFFFFu);
                const float diff = med[c] - (float)xs[c];
                const float a = UNIT ? diff : wi[c] * diff;
                float term = a * a;
                if constexpr (!POS) term = sg[c] * term;
                d = c == 0 ? term : d + term;  // 0 + term0 == term0 up to the sign of zero
            }
            const uint32_t o = d >= tsq ? 1u : 0u;  // pads are masked out after the loop
            if (j < 32) m_lo |= o << j;
            else m_hi |= o << (j - 32);
            if constexpr (MAXT || SUMS) {
                const bool ov = o && (t0 + j < N);
                if constexpr (MAXT) {
                    if (ov && d > lmax) { lmax = d; lmaxi = t0 + j; }
                }
                if constexpr (SUMS) {
#pragma unroll
                    for (int c = 0; c < C; ++c) sum_out[c] += ov ? xs[c] : 0u;
                }
            }
        }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (may_outlier) {
        if (want_max) {
            if (want_sums) classify(F_{}, F_{}, T_{}, T_{});
            else if (unit) classify(T_{}, T_{}, T_{}, F_{});
            else if (pos) classify(F_{}, T_{}, T_{}, F_{});
            else classify(F_{}, F_{}, T_{}, F_{});
        } else {
            if (want_sums) classify(F_{}, F_{}, F_{}, T_{});
            else if (unit) classify(T_{}, T_{}, F_{}, F_{});
            else if (pos) classify(F_{}, T_{}, F_{}, F_{});
            else classify(F_{}, F_{}, F_{}, F_{});
        }
    }
    if (want_sums) {
        // all-frame sums through v_sad_u16 against zero; pads (MAXV) are taken back out
        const uint32_t npad = t0 + F > N ? min(t0 + F - N, (uint32_t)F) : 0u;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uint32_t acc = 0;
#pragma unroll
            for (int i = 0; i < NW; ++i) acc = __builtin_amdgcn_sad_u16(raw[c][i], 0u, acc);
            sum_all[c] = acc - npad * MAXV;
        }
    }

    // ---- drop pads from the masks, then group reductions
    uint32_t v_lo, v_hi;
    {
        const int32_t nv = min((int32_t)N - (int32_t)t0, (int32_t)F);  // valid frames in this block
        v_lo = nv >= 32 ? 0xFFFFFFFFu : (nv <= 0 ? 0u : (1u << nv) - 1u);
        v_hi = nv >= 64 ? 0xFFFFFFFFu : (nv <= 32 ? 0u : (1u << (nv - 32)) - 1u);
    }
    m_lo &= v_lo;
    m_hi &= v_hi;
    const uint32_t num = (uint32_t)grp_sum<G>((int32_t)(__builtin_popcount(m_lo) + __builtin_popcount(m_hi)));
    uint32_t max_index = 0;
    float max_d = 0.0f;
    if (want_max) {
        max_d = __builtin_bit_cast(float, grp_max_u32<G>(__builtin_bit_cast(uint32_t, lmax)));  // d >= 0
        const uint32_t cand = (lmaxi != NONE && lmax == max_d && max_d > 0.0f) ? lmaxi : NONE;
        max_index = grp_min_u32<G>(cand);
        if (max_index == NONE) max_index = 0;
    }
    if (want_sums) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            sum_all[c] = (uint32_t)grp_sum<G>((int32_t)sum_all[c]);
            sum_out[c] = (uint32_t)grp_sum<G>((int32_t)sum_out[c]);
        }
    }
    const uint32_t lane_first = m_lo ? t0 + __builtin_ctz(m_lo) : (m_hi ? t0 + 32 + __builtin_ctz(m_hi) : NONE);
    const uint32_t lane_last = m_hi ? t0 + 63 - __builtin_clz(m_hi) : (m_lo ? t0 + 31 - __builtin_clz(m_lo) : 0u);
    const uint32_t first_idx = grp_min_u32<G>(lane_first);
    const uint32_t last_idx = grp_max_u32<G>(lane_last);
    // first valid non-outlier frame (background "first")
    const uint32_t i_lo = v_lo & ~m_lo, i_hi = v_hi & ~m_hi;
    const uint32_t lane_non = i_lo ? t0 + __builtin_ctz(i_lo) : (i_hi ? t0 + 32 + __builtin_ctz(i_hi) : NONE);
    const uint32_t first_non = grp_min_u32<G>(lane_non);
    // every lane of the group learns the group's masks (for in-order outlier walks)
    const uint32_t lane = __lane_id();
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    uint32_t gm[G][2];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        gm[q][0] = (uint32_t)__shfl((int)m_lo, (int)(gbase + q));
        gm[q][1] = (uint32_t)__shfl((int)m_hi, (int)(gbase + q));
    }

    if (g != 0 || !px_ok) return;

    // ---- finish one pixel: background + outlier selection (src/chrono.rs:290-493)
    constexpr uint32_t maxv = MAXV;
    FastPixel<B, C> P_{k, s_off, pix};
    auto dist_of = [&](uint32_t t, uint32_t (&s)[4]) -> float {
        P_.sample(t, s);
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
        for (int c = 0; c < C; ++c) mean[c] = (float)sum_all[c] / (float)N;
        if (num == 1) {
            P_.sample(first_idx, s);
#pragma unroll
            for (int c = 0; c < C; ++c)
                pixel[c] = sat_cast(roundf(mean[c] * ((float)N / (float)(N - 1)) - (float)s[c] / (float)N), maxv);
        } else if (num > 1) {
#pragma unroll
            for (int c = 0; c < C; ++c)
                pixel[c] = sat_cast(roundf(mean[c] * ((float)N / (float)(N - num)) - (float)sum_out[c] / (float)N), maxv);
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
                bg_index = rng_range(k.seed, pix_id, 0, N);
                warning = num == N;
            } else {
                uint32_t pos = rng_range(k.seed, pix_id, 0, N - num);
                uint32_t j = num;
#pragma unroll
                for (int q = G - 1; q >= 0; --q) {
#pragma unroll
                    for (int w = 1; w >= 0; --w) {
                        uint32_t m = gm[q][w];
                        while (m) {
                            const uint32_t b = 31 - __builtin_clz(m);
                            m &= ~(1u << b);
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
        if (bg_index == 0) {  // frame 0 is this lane's first register half
#pragma unroll
            for (int c = 0; c < C; ++c) s[c] = raw[c][0] & 0xFFFFu;
        } else {
            P_.sample(bg_index, s);
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
            const bool fwd = k.outlier == 4;
#pragma unroll
            for (int qq = 0; qq < G; ++qq) {
                const int q = fwd ? qq : G - 1 - qq;
#pragma unroll
                for (int ww = 0; ww < 2; ++ww) {
                    const int w = fwd ? ww : 1 - ww;
                    uint32_t m = gm[q][w];
                    while (m) {
                        const uint32_t b = fwd ? (uint32_t)__builtin_ctz(m) : 31 - __builtin_clz(m);
                        m &= ~(1u << b);
                        const uint32_t t = (uint32_t)q * F + (uint32_t)w * 32 + b;
                        const float d = dist_of(t, s);
                        const float fade = fade_pick(k.fade, (int32_t)t, (int32_t)N, k.frame_offset);
                        const float blend = fade * blend_value(k.th, sqrtf(d));
                        blend_into_f32<C>(pix_new, s, blend);
                        blend_inv *= 1.0f - blend;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < C; ++c) pixel[c] = sat_cast(roundf(pix_new[c]), maxv);
            blend_byte = sat_cast(roundf((1.0f - blend_inv) * 255.0f), 255u);
        } else {
            uint32_t sel = 0;
            float dist;
            if (k.outlier == 3) {
                float mean_dist = 0.0f;
#pragma unroll
                for (int q = 0; q < G; ++q) {
#pragma unroll
                    for (int w = 0; w < 2; ++w) {
                        uint32_t m = gm[q][w];
                        while (m) {
                            const uint32_t b = __builtin_ctz(m);
                            m &= m - 1;
                            const uint32_t t = (uint32_t)q * F + (uint32_t)w * 32 + b;
                            uint32_t tmp[4];
                            mean_dist += sqrtf(dist_of(t, tmp));
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < C; ++c) s[c] = sat_cast(roundf((float)sum_out[c] / (float)num), maxv);
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

template <int B, int C, int G, int F, bool ABS>
__global__ void __launch_bounds__(256) outlier_fast_kernel(OutlierK k) {
    __shared__ uint64_t s_off[G * F];
    for (uint32_t i = threadIdx.x; i < k.n; i += blockDim.x) s_off[i] = k.pos_off[i];
    __syncthreads();
    outlier_tile<B, C, G, F, ABS>(k, blockIdx.x, blockIdx.y, s_off);
}

}  // namespace cp
