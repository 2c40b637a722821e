// k_outlier_u8.h -- register-resident outlier composite for 8-bit stacks of any channel count,
// with up to 32 lanes per pixel group (stacks of up to 1280 frames) and an optional separate
// block of the `--sample` frames.
//
// Layout.  A 256-lane workgroup covers P = X * 256 / G consecutive output pixels of one row;
// the G adjacent lanes of a pixel group each hold a contiguous block of F frames of the
// group's X pixels (C channels each) as packed bytes, y[e * C + c][i] byte j = channel c of
// pixel e of frame g * F + 4 i + j, so a lane's F frames take X * C * F / 4 VGPRs and the
// rank searches (byte_select_median, byte_select_quartiles) read them as they lie.  Frames past N are padded with 255,
// which never changes a rank below N.
//
// Loads.  The memory pipeline spends about the same time on a wave's load instruction
// whatever its width, so a lane reads each frame's X * C bytes with one dword-aligned load
// (8 or 16 bytes from the dword at or below them) and v_alignbyte drops the bytes before
// them; with X = 4 a frame is 12 useful bytes per lane (RGB).  The loads may read up to 16
// bytes past a tile row's last sample; the launcher checks that those bytes lie inside the
// stack allocation (u8_tile_fits).
//
// `--sample` (src/chrono.rs:151-163, 216-255): the medians and quartiles come from a subset
// of the frame positions while the distances run over all of them.  With FS > 0 every lane
// also loads an FS-frame block of the sample positions (offsets in samp_off) and the ranks
// are searched there (padded with 255 past the subset); the classification still covers the
// F-frame block of all positions.
//
// After the ranks each lane classifies its frames in the reference's f32 order
// (src/chrono.rs:257-288), leaves its outlier bits, largest distance and sums in LDS (the
// 64-frame sorting tile also its outliers' distances, for the average mode), and the
// per-pixel finish (finish_pixel, shared with the 16-bit kernel) selects the background and
// blends (src/chrono.rs:290-493).
#pragma once
#include <utility>

#include "k_outlier_fast.h"

namespace cp {

// v_perm_b32 selector: byte i of the result is byte s_i of {hi, lo} (0-3: lo, 4-7: hi).
constexpr uint32_t psel(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3) {
    return s0 | (s1 << 8) | (s2 << 16) | (s3 << 24);
}

// NW words of bytes from byte address a: word h holds bytes a + 4h .. a + 4h + 3.  They come
// from the NQ dwords at or below a, realigned by a's offset in its dword: NQ = NW + 1 in
// general, NQ = NW when a is dword-aligned (4-channel samples), NQ = NW = 1 when the bytes
// needed lie in a's dword (one pixel of 1 or 2 channels).
template <int NQ, int NW>
__device__ __forceinline__ void load_realigned(const uint8_t* a, uint32_t (&w)[NW]) {
    const uintptr_t ad = (uintptr_t)a;
    const uint32_t sft = (uint32_t)ad & 3u;
    const uintptr_t al = ad & ~(uintptr_t)3;
    uint32_t q[NQ];
    if constexpr (NQ == 1) {
        q[0] = *reinterpret_cast<const uint32_t*>(al);
    } else if constexpr (NQ == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(al);
        q[0] = v.x;
        q[1] = v.y;
    } else if constexpr (NQ == 3) {
        typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
        const u32x3 v = *reinterpret_cast<const u32x3*>(al);
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
    } else {
        static_assert(NQ == 4, "at most 16 bytes per load");
        const uint4 v = *reinterpret_cast<const uint4*>(al);
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
        q[3] = v.w;
    }
    if constexpr (NQ == NW + 1) {
#pragma unroll
        for (int h = 0; h < NW; ++h) w[h] = __builtin_amdgcn_alignbyte(q[h + 1], q[h], sft);
    } else if constexpr (NQ == 1) {
        w[0] = __builtin_amdgcn_alignbyte(0u, q[0], sft);
    } else {
        static_assert(NQ == NW, "dword-aligned words");
#pragma unroll
        for (int h = 0; h < NW; ++h) w[h] = q[h];
    }
}

// 4x4 byte transpose: out[k] byte f = byte k of in[f].
__device__ __forceinline__ void transpose4(const uint32_t (&in)[4], uint32_t (&out)[4]) {
    const uint32_t a_lo = __builtin_amdgcn_perm(in[1], in[0], psel(0, 4, 1, 5));  // f0b0 f1b0 f0b1 f1b1
    const uint32_t a_hi = __builtin_amdgcn_perm(in[1], in[0], psel(2, 6, 3, 7));  // f0b2 f1b2 f0b3 f1b3
    const uint32_t b_lo = __builtin_amdgcn_perm(in[3], in[2], psel(0, 4, 1, 5));
    const uint32_t b_hi = __builtin_amdgcn_perm(in[3], in[2], psel(2, 6, 3, 7));
    out[0] = __builtin_amdgcn_perm(b_lo, a_lo, psel(0, 1, 4, 5));
    out[1] = __builtin_amdgcn_perm(b_lo, a_lo, psel(2, 3, 6, 7));
    out[2] = __builtin_amdgcn_perm(b_hi, a_hi, psel(0, 1, 4, 5));
    out[3] = __builtin_amdgcn_perm(b_hi, a_hi, psel(2, 3, 6, 7));
}

// Loads 4 frames (frame f at base + off[f]) of the lane's X pixels and transposes them into
// v[e * C + c] = the 4 frames' bytes of pixel e, channel c.
template <int C, int X>
__device__ __forceinline__ void load_quad_u8(const uint8_t* __restrict__ base, const uint64_t* off,
                                             uint32_t (&v)[X * C]) {
    constexpr int NB = X * C;         // bytes per frame
    constexpr int NW = (NB + 3) / 4;  // words per frame
    constexpr int NQ = C == 4 ? NW : (X == 1 && C <= 2 ? 1 : NW + 1);  // dwords loaded
    uint32_t w[4][NW];
#pragma unroll
    for (int f = 0; f < 4; ++f) load_realigned<NQ, NW>(base + off[f], w[f]);
#pragma unroll
    for (int h = 0; h < NW; ++h) {
        const uint32_t in[4] = {w[0][h], w[1][h], w[2][h], w[3][h]};
        uint32_t out[4];
        transpose4(in, out);
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (4 * h + b < NB) v[4 * h + b] = out[b];
    }
}

// A lane's block of F frames (offsets off[0..F)) as packed bytes; frames from `valid` on are
// set to 255 (valid is grid-uniform in the tile's block layout: N or the sample count).
template <int C, int X, int F>
__device__ __forceinline__ void load_block_u8(const uint8_t* __restrict__ base, const uint64_t* off,
                                              uint32_t (&y)[X * C][F / 4], uint32_t t0, uint32_t valid, bool pad) {
#pragma unroll
    for (int i = 0; i < F / 4; ++i) {
        uint32_t v[X * C];
        load_quad_u8<C, X>(base, off + 4 * i, v);
#pragma unroll
        for (int c = 0; c < X * C; ++c) y[c][i] = v[c];
    }
    if (pad) {
#pragma unroll
        for (int i = 0; i < F / 4; ++i) {
            const int32_t rem = (int32_t)valid - (int32_t)(t0 + 4 * i);  // valid frames in this dword
            const uint32_t m = rem >= 4 ? 0u : (rem <= 0 ? 0xFFFFFFFFu : (0xFFFFFFFFu << (8 * rem)));
#pragma unroll
            for (int c = 0; c < X * C; ++c) y[c][i] |= m;
        }
    }
}

// min{x > v} over the group's packed bytes (exists by the caller's count); the bytes are
// widened to 16-bit halves, where values at or below v wrap to the top.
template <int G, int ND>
__device__ __forceinline__ uint32_t min_greater_u8(const uint32_t (&y)[ND], uint32_t v) {
    const us2 V = {(unsigned short)(v + 1), (unsigned short)(v + 1)};
    us2 m = {0xFFFF, 0xFFFF};
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        const uint32_t lo = y[i] & 0x00FF00FFu, hi = (y[i] >> 8) & 0x00FF00FFu;
        m = __builtin_elementwise_min(m, as_us2(lo) - V);
        m = __builtin_elementwise_min(m, as_us2(hi) - V);
    }
    const uint32_t mm = min((uint32_t)m.x, (uint32_t)m.y);
    return grp_min_u32<G>(mm) + v + 1;
}

// Values at ranks (r0, r1), r1 in {r0, r0 + 1}, for the median and, for relative thresholds,
// both quartiles of every channel: the median from a 16-value window around an estimate
// (byte_select_median; real = the lane block's frames below the count the ranks are over), the
// quartiles from windows next to it, the C chains of each search interleaved.
template <int G, int ND, int C, bool ABS>
__device__ __forceinline__ void ranks_u8(const uint32_t (&y)[C][ND], const RankPlan& plan, int32_t ntot,
                                         uint32_t real, uint32_t (&m0)[C], uint32_t (&m1)[C], uint32_t (&a0)[C],
                                         uint32_t (&a1)[C], uint32_t (&b0)[C], uint32_t (&b1)[C]) {
    constexpr int R = ABS ? 1 : 3;
    const uint32_t kk3[3] = {plan.m0, plan.a0, plan.b0};
    uint32_t sel[C][R], cle[C][R];
    if constexpr (ABS) {
        uint32_t o[C], n[C];
        byte_select_median<G, ND, C>(y, plan.m0, ntot, real, o, n);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            sel[c][0] = o[c];
            cle[c][0] = n[c];
        }
    } else {
        byte_select_quartiles<G, ND, C>(y, kk3, ntot, real, sel, cle);
    }
    auto nb = [&](int c, uint32_t v0, uint32_t cnt, uint32_t r0, uint32_t r1) -> uint32_t {
        if (r1 == r0 || cnt >= r1 + 1) return v0;
        return min_greater_u8<G, ND>(y[c], v0);
    };
#pragma unroll
    for (int c = 0; c < C; ++c) {
        m0[c] = sel[c][0];
        m1[c] = nb(c, m0[c], cle[c][0], plan.m0, plan.m1);
        if constexpr (!ABS) {
            a0[c] = sel[c][R > 1 ? 1 : 0];
            a1[c] = nb(c, a0[c], cle[c][R > 1 ? 1 : 0], plan.a0, plan.a1);
            b0[c] = sel[c][R > 2 ? 2 : 0];
            b1[c] = nb(c, b0[c], cle[c][R > 2 ? 2 : 0], plan.b0, plan.b1);
        } else {
            a0[c] = a1[c] = b0[c] = b1[c] = 0;
        }
    }
}

// ---- sorting-network ranks (one lane per pixel)
//
// With G = 1 a lane holds every frame of its pixel, so the ranks can come from sorting instead
// of the byte search: Batcher's odd-even merge sort of the F frames (F a power of two),
// two channels at a time in the 16-bit halves of a dword (v_pk_min_u16 / v_pk_max_u16 per
// comparator), then the needed ranks picked from the sorted registers; an odd channel (the
// third of RGB, or the only one) takes the byte search instead.  Any exact sort gives the
// reference's values (src/chrono.rs:228-255 sorts too); pads (255) sort to the top.  It does
// every rank of two channels in one pass, which pays off against six byte searches for
// relative thresholds; the autotuner measures both per frame count.
struct CxPair {
    int a, b;
};

template <int N>
constexpr int oems_count() {
    int c = 0;
    for (int p = 1; p < N; p <<= 1)
        for (int k = p; k >= 1; k >>= 1)
            for (int j = k % p; j + k < N; j += 2 * k)
                for (int i = 0; i < k && i + j + k < N; ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) ++c;
    return c;
}

template <int N>
struct OemsNet {
    static constexpr int M = oems_count<N>();
    CxPair p[M];
    constexpr OemsNet() : p{} {
        int c = 0;
        for (int q = 1; q < N; q <<= 1)
            for (int k = q; k >= 1; k >>= 1)
                for (int j = k % q; j + k < N; j += 2 * k)
                    for (int i = 0; i < k && i + j + k < N; ++i)
                        if ((i + j) / (2 * q) == (i + j + k) / (2 * q)) p[c++] = CxPair{i + j, i + j + k};
    }
};

template <int N>
constexpr OemsNet<N> kOemsNet{};

template <int A, int B, int N>
__device__ __forceinline__ void cx_pair(us2 (&v)[N]) {
    const us2 lo = __builtin_elementwise_min(v[A], v[B]), hi = __builtin_elementwise_max(v[A], v[B]);
    v[A] = lo;
    v[B] = hi;
}

template <int N, size_t... Is>
__device__ __forceinline__ void oems_apply(us2 (&v)[N], std::index_sequence<Is...>) {
    (cx_pair<kOemsNet<N>.p[Is].a, kOemsNet<N>.p[Is].b, N>(v), ...);
}

// v[r] for a wave-uniform rank r: the registers as vectors of up to 32 read at a dynamic
// index, which compiles to register-relative moves (a chain of selects is turned into a copy
// of v through scratch memory and an indexed load, and so is one read of a longer vector).
template <int N>
__device__ __forceinline__ us2 pick_rank(const us2 (&v)[N], uint32_t r) {
    constexpr int W = N < 32 ? N : 32;
    typedef uint32_t vecW __attribute__((ext_vector_type(W)));
    const uint32_t ru = __builtin_amdgcn_readfirstlane(r);
    uint32_t s = 0;
#pragma unroll
    for (int b = 0; b < N; b += W) {
        vecW w;
#pragma unroll
        for (int j = 0; j < W; ++j) w[j] = __builtin_bit_cast(uint32_t, v[b + j]);
        const uint32_t e = w[ru % W];
        s = ru / W == (uint32_t)(b / W) ? e : s;
    }
    return as_us2(s);
}

// The rank values of ranks_u8 from a sort of the lane's F frames (G = 1).
template <int F, int C, bool ABS>
__device__ __forceinline__ void ranks_sort(const uint32_t (&y)[C][F / 4], const RankPlan& plan, uint32_t real,
                                           uint32_t (&m0)[C],
                                           uint32_t (&m1)[C], uint32_t (&a0)[C], uint32_t (&a1)[C],
                                           uint32_t (&b0)[C], uint32_t (&b1)[C]) {
    static_assert((F & (F - 1)) == 0, "the network sorts a power of two");
    constexpr int NP = C / 2;  // channel pairs sorted together
    if constexpr (C % 2 == 1) {
        // the odd channel alone: a network pass on one channel costs as much as on two, more
        // than the byte search of one channel's ranks
        uint32_t yl[1][F / 4], o0[1], o1[1], oa0[1], oa1[1], ob0[1], ob1[1];
#pragma unroll
        for (int i = 0; i < F / 4; ++i) yl[0][i] = y[C - 1][i];
        ranks_u8<1, F / 4, 1, ABS>(yl, plan, F, real, o0, o1, oa0, oa1, ob0, ob1);
        m0[C - 1] = o0[0];
        m1[C - 1] = o1[0];
        a0[C - 1] = oa0[0];
        a1[C - 1] = oa1[0];
        b0[C - 1] = ob0[0];
        b1[C - 1] = ob1[0];
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int c0 = 2 * q, c1 = 2 * q + 1;
        us2 v[F];
#pragma unroll
        for (int j = 0; j < F; ++j)  // byte j&3 of each channel's dword into the two halves
            v[j] = as_us2(__builtin_amdgcn_perm(y[c1][j >> 2], y[c0][j >> 2],
                                                psel(j & 3, 12, 4 + (j & 3), 12)));
        oems_apply<F>(v, std::make_index_sequence<kOemsNet<F>.M>{});
        auto put = [&](uint32_t (&dst)[C], uint32_t r) {
            const us2 s = pick_rank<F>(v, r);
            dst[c0] = s.x;
            dst[c1] = s.y;
        };
        put(m0, plan.m0);
        put(m1, plan.m1);
        if constexpr (!ABS) {
            put(a0, plan.a0);
            put(a1, plan.a1);
            put(b0, plan.b0);
            put(b1, plan.b1);
        }
    }
    if constexpr (ABS) {
#pragma unroll
        for (int c = 0; c < C; ++c) a0[c] = a1[c] = b0[c] = b1[c] = 0;
    }
}

// Whether a tile keeps its outliers' squared distances in LDS for the average mode's finish
// (finish_pixel's odist): one lane per pixel, so the lane that classifies a pixel is the lane
// that finishes it, and 64-frame lanes, whose pixels have outliers enough for the re-reads
// to cost more than the stores (the LDS it takes, 16 KB, would cut the occupancy of the
// shorter one-lane tiles).
template <int G, int F, int X, int FS, bool SORT>
constexpr bool keeps_odist() {
    return SORT && G == 1 && X == 1 && F == 64 && FS == 0;
}

// One tile: P = X * 256 / G pixels of row y from pixel bx * P.  sh.off holds the tile's frame
// offsets (all positions) and soff the sample positions' offsets when FS > 0; every lane of
// the workgroup calls this (it synchronises once).  lane: threadIdx.x.  odist: (kOdistSlots + 1)
// * P floats of LDS when keeps_odist(), else unused.
template <int C, int G, int F, int X, int FS, bool ABS, bool SORT = false>
__device__ __forceinline__ void outlier_u8_tile(const OutlierK& k, uint32_t bx, uint32_t y,
                                                TileShared<C, G, F, X>& sh, const uint64_t* soff, uint32_t lane,
                                                float* odist = nullptr) {
    static_assert(F % 8 == 0 && F <= 96 && G <= 32 && FS % 4 == 0 && (X == 1 || X == 4), "tile shape");
    static_assert(!SORT || (G == 1 && X == 1 && FS == 0), "sorting-network ranks: one lane per pixel");
    using TS = TileShared<C, G, F, X>;
    constexpr int P = TS::P;
    constexpr int MW = TS::MW;
    constexpr int ND = F / 4;
    constexpr int XC = X * C;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    const uint32_t N = k.n;
    const uint32_t g = lane % G;
    const uint32_t t0 = g * F;
    {
        const uint32_t pl = lane / G;
        const uint32_t x0 = bx * P + pl * X;
        const uint64_t pix = ((uint64_t)y * k.stack_w + (x0 < k.out_w ? x0 : 0u)) * C;
        const uint8_t* base = k.stack + pix;

        uint32_t yv[XC][ND];
        load_block_u8<C, X, F>(base, sh.off + g * TS::OS, yv, t0, N, N < (uint32_t)(G * F));

        // ---- medians / inverse IQR over the sample set, every pixel and channel
        const RankPlan plan = rank_plan_core(k.n_sample, ABS);
        uint32_t m0[XC], m1[XC], a0[XC], a1[XC], b0[XC], b1[XC];
        if constexpr (SORT) {
            ranks_sort<F, XC, ABS>(yv, plan, min(N, (uint32_t)F), m0, m1, a0, a1, b0, b1);
        } else if constexpr (FS > 0) {
            constexpr int NDS = FS / 4;
            uint32_t ys[XC][NDS];
            load_block_u8<C, X, FS>(base, soff + g * (FS + 2), ys, g * FS, k.n_sample,
                                    k.n_sample < (uint32_t)(G * FS));
            const uint32_t ts = g * FS, ns = k.n_sample;
            ranks_u8<G, NDS, XC, ABS>(ys, plan, G * FS, ns > ts ? min(ns - ts, (uint32_t)FS) : 0u, m0, m1, a0, a1,
                                      b0, b1);
        } else {
            ranks_u8<G, ND, XC, ABS>(yv, plan, G * F, N > t0 ? min(N - t0, (uint32_t)F) : 0u, m0, m1, a0, a1, b0,
                                     b1);
        }
        float med[X][4], iqr_inv[X][4];
#pragma unroll
        for (int e = 0; e < X; ++e) {
#pragma unroll
            for (int c = 0; c < 4; ++c) med[e][c] = iqr_inv[e][c] = 0.0f;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float mc, ic;
                const int q = e * C + c;
                stats_from_values(plan, m0[q], m1[q], a0[q], a1[q], b0[q], b1[q], ABS, &mc, &ic);
                med[e][c] = k.w[c] == 0.0f ? 0.0f : mc;  // channels without weight keep zero statistics
                iqr_inv[e][c] = k.w[c] == 0.0f ? 0.0f : ic;
            }
        }

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
        // threshold, so `d >= lmax` also implies that the frame is an outlier.  With
        // keeps_odist() the average mode's outlier distances go to LDS in that order.
        constexpr bool ODIST = keeps_odist<G, F, X, FS, SORT>();
        const bool store_d = ODIST && k.outlier == 3;
        uint32_t nst = 0;
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
                    xs[c] = (yv[E * C + c][j >> 2] >> (8 * (j & 3))) & 0xFFu;
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
                    if (in && d >= lmax[E]) {
                        lmax[E] = d;
                        lmaxi[E] = t0 + (uint32_t)j;
                    }
                }
                if constexpr (SUMS && CHECK) {
#pragma unroll
                    for (int c = 0; c < C; ++c) sum_out[E][c] += (o && in) ? xs[c] : 0u;
                }
                if constexpr (ODIST && SUMS) {
                    // no branch: frames that are not stored go to the spare slot kOdistSlots
                    const bool put = store_d && o && in;
                    odist[(put ? min(nst, (uint32_t)kOdistSlots) : (uint32_t)kOdistSlots) * (uint32_t)P + lane] = d;
                    nst += put ? 1u : 0u;
                }
            }
            if constexpr (SUMS && !CHECK) {
                // every frame is real: the outliers' sums from the mask, four frames per
                // v_dot4_u32_u8 (the mask's four bits spread to one 0/1 byte per frame)
#pragma unroll
                for (int i = 0; i < ND; ++i) {
                    const uint32_t nib = (m[E][i >> 3] >> (4 * (i & 7))) & 0xFu;
                    const uint32_t sel = (nib * 0x00204081u) & 0x01010101u;
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        sum_out[E][c] = __builtin_amdgcn_udot4(yv[E * C + c][i], sel, sum_out[E][c], false);
                }
            }
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        StaticFor<0, X>::run([&](auto e_t) {
            constexpr int E = decltype(e_t)::value;
            // Bound check (lanes with few frames): with non-negative weights every frame's f32
            // distance is at most the same f32 expression on the per-channel maximum deviation
            // from the median (correct rounding is monotone), so below the threshold no frame
            // of this lane is an outlier and the per-frame pass is skipped.
            bool may_outlier = true;
            if constexpr (F <= 32) {
                if (pos && full_tile) {
                    float bound = 0.0f;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        us2 mn = {0xFFFF, 0xFFFF}, mx = {0, 0};
#pragma unroll
                        for (int i = 0; i < ND; ++i) {
                            const uint32_t v = yv[E * C + c][i];
                            const us2 lo = as_us2(v & 0x00FF00FFu), hi = as_us2((v >> 8) & 0x00FF00FFu);
                            mn = __builtin_elementwise_min(mn, __builtin_elementwise_min(lo, hi));
                            mx = __builtin_elementwise_max(mx, __builtin_elementwise_max(lo, hi));
                        }
                        const float lo = (float)min((uint32_t)mn.x, (uint32_t)mn.y);
                        const float hi = (float)max((uint32_t)mx.x, (uint32_t)mx.y);
                        const float dev = fmaxf(med[E][c] - lo, hi - med[E][c]);
                        const float a = unit ? dev : wi[E][c] * dev;
                        bound = c == 0 ? a * a : bound + a * a;
                    }
                    may_outlier = bound >= tsq;
                }
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
            // all-frame sums through v_sad_u8 against zero; pads (255) are taken back out
            const uint32_t npad = t0 + F > N ? min(t0 + F - N, (uint32_t)F) : 0u;
#pragma unroll
            for (int e = 0; e < X; ++e)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int i = 0; i < ND; ++i) acc = __builtin_amdgcn_sad_u8(yv[e * C + c][i], 0u, acc);
                    sum_all[e][c] = (uint32_t)grp_sum<G>((int32_t)(acc - npad * 255u));
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
                    if (want_s0) sh.s0[c][p] = yv[e * C + c][0] & 0xFFu;
                    if (want_sums) {
                        sh.sum_all[c][p] = sum_all[e][c];
                        sh.sum_out[c][p] = sum_out[e][c];
                    }
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t p = lane; p < (uint32_t)P; p += 256)
        finish_pixel<1, C, G, F, X, ABS>(k, bx, y, p, sh, keeps_odist<G, F, X, FS, SORT>() ? odist : nullptr);
}

// VGPRs the frame blocks take, for the occupancy hint: X * C * (F + FS) / 4, plus F for the
// pairs a sorting network works on, plus working registers.
constexpr int waves_hint_u8(int C, int F, int X, int FS, bool sort) {
    const int w = 448 / (X * C * (F + FS) / 4 + (sort ? F : 0) + 64);
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <int C, int G, int F, int X, int FS, bool ABS, bool SORT = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint_u8(C, F, X, FS, SORT))))
outlier_u8_kernel(OutlierK k, uint32_t tiles_x) {
    __shared__ TileShared<C, G, F, X> sh;
    __shared__ uint64_t soff[FS > 0 ? G * (FS + 2) : 1];
    __shared__ float odist[keeps_odist<G, F, X, FS, SORT>() ? (kOdistSlots + 1) * TileShared<C, G, F, X>::P : 1];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    load_offsets<1, F>(sh.off, G * F, k.pos_off, k.n);
    if constexpr (FS > 0) load_offsets<1, FS>(soff, G * FS, k.samp_off, k.n_sample);
    __syncthreads();
    outlier_u8_tile<C, G, F, X, FS, ABS, SORT>(k, tile % tiles_x, tile / tiles_x, sh, soff, threadIdx.x, odist);
}

// Whether the packed-byte tiles may run over this composite: every lane's load of a frame
// (X * C bytes read with up to 8 bytes around them from the dword at or below them) stays
// inside the stack allocation, and 4-channel stacks start on a dword boundary.
inline bool u8_tile_fits(const void* stack, uint64_t alloc_bytes, uint32_t channels, int X, const OutlierK& k,
                         uint64_t max_pos) {
    if (channels == 4 && ((uintptr_t)stack & 3u)) return false;
    const uint64_t x_last = (uint64_t)((k.out_w - 1) / X) * X;
    const uint64_t first = max_pos + ((uint64_t)(k.out_h - 1) * k.stack_w + x_last) * channels;
    return first + (uint64_t)X * channels + 8 <= alloc_bytes;
}

}  // namespace cp
