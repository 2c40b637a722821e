/*
 * chrono_oracle.c -- CPU restatement of chrono-photo v0.6.5's compositing path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, smoke() and bench.py's
 * cpu_baseline leg load this library; the product never does.  See chrono_oracle.h for
 * the parity pins of each function.
 *
 * Build: gcc -O3 -ffp-contract=off -fno-fast-math (-fopenmp for the threaded baseline).
 *
 * Generalisation beyond the reference (which is 8-bit only, src/slicer.rs:141-143):
 * bytes_per_sample = 2 reads u16 samples, absolute thresholds scale by 65535 instead of
 * 255, output pixels saturate to u16, the blend mask stays u8 and the shake SSD is
 * accumulated in int64 (65535^2 overflows i32).
 */
#include "chrono_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- helpers */

/* Rust `f as u8` / `f as u16`: NaN -> 0, saturate, truncate toward zero. */
static inline uint32_t sat_cast(float f, uint32_t maxv) {
    if (!(f > 0.0f)) return 0u;
    if (f >= (float)maxv) return maxv;
    return (uint32_t)f;
}

/* Rust f32::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives. */
static inline float signum_f(float w) { return signbit(w) ? -1.0f : 1.0f; }

static inline uint32_t sample_at(const or_view* v, uint32_t pos, uint32_t x, uint32_t y,
                                 uint32_t ch) {
    uint32_t slot = v->slots ? v->slots[pos] : pos;
    uint32_t xx = x, yy = y;
    if (v->crop_xy) {
        xx += (uint32_t)v->crop_xy[2 * pos];
        yy += (uint32_t)v->crop_xy[2 * pos + 1];
    }
    size_t idx = (((size_t)slot * v->height + yy) * v->width + xx) * v->channels + ch;
    if (v->bytes_per_sample == 1) return ((const uint8_t*)v->data)[idx];
    return ((const uint16_t*)v->data)[idx];
}

static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint32_t or_rng_range(uint64_t seed, uint64_t pixel, uint32_t draw, uint32_t n) {
    uint64_t h = splitmix64(seed ^ splitmix64(pixel * 4ull + (uint64_t)draw));
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

/* ------------------------------------------------------ options.rs / color.rs */

/* Threshold::new, src/options.rs:197-214 (range 255; 65535 for u16 samples). */
void or_threshold_new(int32_t absolute, float min, float max, float range, or_threshold* out) {
    out->absolute = absolute;
    if (absolute) {
        out->min = min * range;
        out->max = max * range;
        out->scale = 1.0f / ((max - min) * range);
    } else {
        out->min = min;
        out->max = max;
        out->scale = 1.0f / (max - min);
    }
}

/* Threshold::blend_value, src/options.rs:224-231. */
float or_blend_value(const or_threshold* t, float dist) {
    if (dist <= t->min) return 0.0f;
    if (dist >= t->max) return 1.0f;
    return (dist - t->min) * t->scale;
}

/* Fade::new, src/options.rs:69-95: linear interpolation between keyframes. */
uint32_t or_fade_table(const int32_t* frames, const float* values, uint32_t nkeys, float* out,
                       uint32_t cap) {
    if (nkeys < 2) return 0;
    int32_t offset = frames[0];
    uint32_t len = (uint32_t)(frames[nkeys - 1] - offset);
    uint32_t idx = 0;
    for (uint32_t i = 0; i < len + 1; ++i) {
        int32_t f1 = frames[idx], f2 = frames[idx + 1];
        float v1 = values[idx], v2 = values[idx + 1];
        int32_t frame = (int32_t)i + offset;
        float val = v1 + (v2 - v1) * (float)(frame - f1) / (float)(f2 - f1);
        if (i < cap) out[i] = val;
        if (frame == f2) idx += 1;
    }
    return len + 1;
}

/* Fade::get, src/options.rs:113-140. */
float or_fade_get(const or_fade* f, int32_t frame) {
    if (f->is_none) return 1.0f;
    int32_t i = frame - f->offset;
    int32_t len = (int32_t)f->len;
    if (i >= 0 && i < len) return f->values[i];
    if (!f->repeat) return i < 0 ? f->values[0] : f->values[len - 1];
    while (i < 0) i += len;
    i = i % len;
    return f->values[i];
}

/* OutlierProcessor::fade src/chrono.rs:496-502 == SimpleProcessor::fade src/simple.rs:170-176. */
float or_fade_pick(const or_fade* f, int32_t frame, int32_t total, int32_t offset) {
    if (f->absolute) return or_fade_get(f, offset + frame);
    return or_fade_get(f, total - frame - 1);
}

/* color::blend_into_u8, src/color.rs:4-16 (u8 samples). */
void or_blend_into_u8(uint8_t* a, const uint8_t* b, uint32_t n, float blend) {
    if (blend <= 0.0f) return;
    if (blend >= 1.0f) {
        for (uint32_t i = 0; i < n; ++i) a[i] = b[i];
        return;
    }
    for (uint32_t i = 0; i < n; ++i) {
        float aa = (float)a[i];
        a[i] = (uint8_t)sat_cast(roundf(aa + ((float)b[i] - aa) * blend), 255u);
    }
}

/* Same as above on u32 containers with a saturation limit (u8 or u16). */
static void blend_into_px(uint32_t* a, const uint32_t* b, uint32_t n, float blend, uint32_t maxv) {
    if (blend <= 0.0f) return;
    if (blend >= 1.0f) {
        for (uint32_t i = 0; i < n; ++i) a[i] = b[i];
        return;
    }
    for (uint32_t i = 0; i < n; ++i) {
        float aa = (float)a[i];
        a[i] = sat_cast(roundf(aa + ((float)b[i] - aa) * blend), maxv);
    }
}

/* color::blend_into_f32_u8, src/color.rs:32-44. */
static void blend_into_f32(float* a, const uint32_t* b, uint32_t n, float blend) {
    if (blend <= 0.0f) return;
    if (blend >= 1.0f) {
        for (uint32_t i = 0; i < n; ++i) a[i] = (float)b[i];
        return;
    }
    for (uint32_t i = 0; i < n; ++i) {
        float aa = a[i];
        a[i] = aa + ((float)b[i] - aa) * blend;
    }
}

/* ------------------------------------------------------------- chrono.rs */

/* OutlierProcessor::median, src/chrono.rs:580-591. */
float or_median(const uint32_t* d, uint32_t len) {
    if ((len + 1) % 2 == 0) return (float)d[(len + 1) / 2 - 1];
    uint32_t idx = (len + 1) / 2;
    return 0.5f * ((float)d[idx - 1] + (float)d[idx]);
}

/* OutlierProcessor::quantile, src/chrono.rs:566-578 (callers guarantee len >= 3). */
float or_quantile(const uint32_t* d, uint32_t len, float q) {
    float pos = (float)(len + 1) * q;
    uint32_t p1 = (uint32_t)pos - 1u;
    float frac = pos - truncf(pos);
    if (frac < 0.001f) return (float)d[p1];
    if (frac > 0.999f) return (float)d[p1 + 1];
    return (1.0f - frac) * (float)d[p1] + frac * (float)d[p1 + 1];
}

/* sort_unstable on integers has a unique result; LSD radix sort by bytes. */
static void sort_u32(uint32_t* a, uint32_t n, uint32_t* tmp, uint32_t bytes) {
    if (n < 24) {
        for (uint32_t i = 1; i < n; ++i) {
            uint32_t v = a[i];
            uint32_t j = i;
            while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; --j; }
            a[j] = v;
        }
        return;
    }
    for (uint32_t pass = 0; pass < bytes; ++pass) {
        uint32_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        uint32_t sh = pass * 8;
        for (uint32_t i = 0; i < n; ++i) cnt[((a[i] >> sh) & 255u) + 1]++;
        for (uint32_t i = 0; i < 256; ++i) cnt[i + 1] += cnt[i];
        for (uint32_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> sh) & 255u]++] = a[i];
        memcpy(a, tmp, n * sizeof(uint32_t));
    }
}

typedef struct {
    uint32_t* values;     /* 4 * S */
    uint32_t* sort_tmp;   /* S */
    uint32_t* out_idx;    /* N */
    float* out_dist;      /* N */
    uint32_t* non_out;    /* N */
} scratch_t;

/* first_excluded, src/chrono.rs:505-530. */
static uint32_t first_excluded(const scratch_t* s, uint32_t samples, uint32_t k, int* warn) {
    if (k == samples) { *warn = 1; return 0; }
    uint32_t e = 0;
    for (uint32_t i = 0; i < samples; ++i) {
        if (e < k && i == s->out_idx[e]) e++;
        else { *warn = 0; return i; }
    }
    *warn = 0;
    return 0; /* unreachable: k < samples guarantees an excluded index */
}

/* sample_excluded, src/chrono.rs:532-555 (swap-based; candidates may hold outliers). */
static uint32_t sample_excluded(scratch_t* s, uint32_t samples, uint32_t k, uint64_t seed,
                                uint64_t pix, uint32_t* draw, int* warn) {
    if (k == samples) {
        *warn = 1;
        return or_rng_range(seed, pix, (*draw)++, samples);
    }
    for (uint32_t i = 0; i < samples; ++i) s->non_out[i] = i;
    uint32_t cand = samples;
    for (uint32_t j = 0; j < k; ++j) {
        uint32_t a = s->out_idx[j], b = cand - 1;
        uint32_t t = s->non_out[a];
        s->non_out[a] = s->non_out[b];
        s->non_out[b] = t;
        cand -= 1;
    }
    uint32_t idx = or_rng_range(seed, pix, (*draw)++, cand);
    *warn = 0;
    return s->non_out[idx];
}

/* calc_pixel, src/chrono.rs:208-494. */
static void calc_pixel(const or_view* v, uint32_t x, uint32_t y, int32_t frame_offset,
                       const uint32_t* sample_pos, uint32_t S, const or_outlier_params* p,
                       scratch_t* s, uint32_t* pixel, uint8_t* blend_out, int* warning,
                       or_pixel_stats* st) {
    const uint32_t C = v->channels;
    const uint32_t N = v->n;
    const uint32_t maxv = v->bytes_per_sample == 1 ? 255u : 65535u;
    const or_threshold* th = &p->threshold;
    const float threshold_sq = th->min * th->min;
    const uint64_t pix_id = (uint64_t)y * v->out_w + x;
    uint32_t draw = 0;

    float median[4] = {0, 0, 0, 0};
    float iqr_inv[4] = {0, 0, 0, 0};

    /* :216-226 gather the sample set per weighted channel */
    for (uint32_t si = 0; si < S; ++si) {
        uint32_t t = sample_pos ? sample_pos[si] : si;
        for (uint32_t ch = 0; ch < C; ++ch)
            if (p->weights[ch] != 0.0f) s->values[ch * S + si] = sample_at(v, t, x, y, ch);
    }
    /* :228-255 medians and inverse IQR */
    for (uint32_t i = 0; i < C; ++i) {
        if (p->weights[i] == 0.0f) continue;
        uint32_t* sl = s->values + i * S;
        sort_u32(sl, S, s->sort_tmp, v->bytes_per_sample);
        if (th->absolute) {
            median[i] = or_median(sl, S);
        } else {
            float q1 = or_quantile(sl, S, 0.25f);
            float med = or_median(sl, S);
            float q3 = or_quantile(sl, S, 0.75f);
            median[i] = med;
            iqr_inv[i] = q3 - q1;
            if (iqr_inv[i] == 0.0f) iqr_inv[i] = 1.0f;
            iqr_inv[i] = 1.0f / iqr_inv[i];
        }
    }

    /* :257-288 distances and classification over all frames */
    uint32_t k = 0;
    float max_dist_sq = 0.0f;
    uint32_t max_index = 0;
    for (uint32_t t = 0; t < N; ++t) {
        float dist_sq = 0.0f;
        for (uint32_t i = 0; i < C; ++i) {
            float w = p->weights[i];
            if (w == 0.0f) continue;
            float diff = median[i] - (float)sample_at(v, t, x, y, i);
            float term;
            if (diff == 0.0f) {
                term = 0.0f;
            } else if (th->absolute) {
                float a = w * diff;
                term = signum_f(w) * (a * a);
            } else {
                float a = w * iqr_inv[i] * diff;
                term = signum_f(w) * (a * a);
            }
            dist_sq += term;
        }
        if (dist_sq >= threshold_sq) {
            s->out_idx[k] = t;
            s->out_dist[k] = dist_sq;
            k++;
            if (dist_sq > max_dist_sq) {
                max_dist_sq = dist_sq;
                max_index = t;
            }
        }
    }

    int has_warning = 0;
    uint32_t bg_index = 0;
    /* :290-375 background */
    switch (p->background) {
    case OR_BG_AVERAGE: {
        float mean[4] = {0, 0, 0, 0};
        for (uint32_t t = 0; t < N; ++t)
            for (uint32_t i = 0; i < C; ++i) mean[i] += (float)sample_at(v, t, x, y, i);
        for (uint32_t i = 0; i < 4; ++i) mean[i] /= (float)N;
        if (k > 0) {
            if (k == 1) {
                uint32_t t0 = s->out_idx[0];
                for (uint32_t ch = 0; ch < C; ++ch)
                    pixel[ch] = sat_cast(roundf(mean[ch] * ((float)N / (float)(N - 1)) -
                                                (float)sample_at(v, t0, x, y, ch) / (float)N),
                                         maxv);
            } else {
                float osum[4] = {0, 0, 0, 0};
                for (uint32_t j = 0; j < k; ++j)
                    for (uint32_t ch = 0; ch < C; ++ch)
                        osum[ch] += (float)sample_at(v, s->out_idx[j], x, y, ch);
                uint32_t non = N - k;
                for (uint32_t ch = 0; ch < C; ++ch)
                    pixel[ch] = sat_cast(
                        roundf(mean[ch] * ((float)N / (float)non) - osum[ch] / (float)N), maxv);
            }
        } else {
            for (uint32_t ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(mean[ch]), maxv);
        }
        break;
    }
    case OR_BG_MEDIAN:
        for (uint32_t ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(median[ch]), maxv);
        break;
    default: { /* First | Random */
        int w = 0;
        if (p->background == OR_BG_FIRST) {
            bg_index = k == 0 ? 0 : first_excluded(s, N, k, &w);
        } else {
            if (k == 0) bg_index = or_rng_range(p->seed, pix_id, draw++, N);
            else bg_index = sample_excluded(s, N, k, p->seed, pix_id, &draw, &w);
        }
        for (uint32_t ch = 0; ch < C; ++ch) pixel[ch] = sample_at(v, bg_index, x, y, ch);
        if (w) has_warning = 1;
        break;
    }
    }

    /* :377-493 outlier selection and blending */
    uint8_t blend_byte = 0;
    if (k > 0) {
        uint32_t smp[4];
        if (k == 1) {
            uint32_t t0 = s->out_idx[0];
            for (uint32_t ch = 0; ch < C; ++ch) smp[ch] = sample_at(v, t0, x, y, ch);
            float fade = or_fade_pick(&p->fade, (int32_t)t0, (int32_t)N, frame_offset);
            float blend = fade * or_blend_value(th, sqrtf(s->out_dist[0]));
            blend_into_px(pixel, smp, C, blend, maxv);
            blend_byte = (uint8_t)sat_cast(roundf(blend * 255.0f), 255u);
        } else if (p->outlier == OR_SEL_FORWARD || p->outlier == OR_SEL_BACKWARD) {
            float pix_new[4] = {0, 0, 0, 0};
            float blend_inv = 1.0f;
            for (uint32_t ch = 0; ch < C; ++ch) pix_new[ch] = (float)pixel[ch];
            for (uint32_t jj = 0; jj < k; ++jj) {
                uint32_t j = p->outlier == OR_SEL_FORWARD ? jj : k - 1 - jj;
                uint32_t t = s->out_idx[j];
                for (uint32_t ch = 0; ch < C; ++ch) smp[ch] = sample_at(v, t, x, y, ch);
                float fade = or_fade_pick(&p->fade, (int32_t)t, (int32_t)N, frame_offset);
                float blend = fade * or_blend_value(th, sqrtf(s->out_dist[j]));
                blend_into_f32(pix_new, smp, C, blend);
                blend_inv *= 1.0f - blend;
            }
            for (uint32_t ch = 0; ch < C; ++ch) pixel[ch] = sat_cast(roundf(pix_new[ch]), maxv);
            blend_byte = (uint8_t)sat_cast(roundf((1.0f - blend_inv) * 255.0f), 255u);
        } else {
            uint32_t sel_idx;
            float dist;
            if (p->outlier == OR_SEL_AVERAGE) {
                /* :446-466 (k > 1 here).  The reference also blends a 4-entry temp sample
                 * into a C-entry pixel; with C < 4 and blend >= 1 that indexes past the
                 * pixel and panics (src/color.rs:8-10).  This restatement writes the C
                 * channels, which is what the reference computes whenever it does not
                 * panic. */
                float mean[4] = {0, 0, 0, 0};
                float mean_dist = 0.0f;
                for (uint32_t j = 0; j < k; ++j) {
                    for (uint32_t ch = 0; ch < C; ++ch)
                        mean[ch] += (float)sample_at(v, s->out_idx[j], x, y, ch);
                    mean_dist += sqrtf(s->out_dist[j]);
                }
                for (uint32_t ch = 0; ch < C; ++ch)
                    smp[ch] = sat_cast(roundf(mean[ch] / (float)k), maxv);
                sel_idx = 0;
                dist = mean_dist / (float)k;
            } else {
                float dsq;
                if (p->outlier == OR_SEL_FIRST) { sel_idx = s->out_idx[0]; dsq = s->out_dist[0]; }
                else if (p->outlier == OR_SEL_LAST) { sel_idx = s->out_idx[k - 1]; dsq = s->out_dist[k - 1]; }
                else { sel_idx = max_index; dsq = max_dist_sq; }
                for (uint32_t ch = 0; ch < C; ++ch) smp[ch] = sample_at(v, sel_idx, x, y, ch);
                dist = sqrtf(dsq);
            }
            float fade = or_fade_pick(&p->fade, (int32_t)sel_idx, (int32_t)N, frame_offset);
            float blend = fade * or_blend_value(th, dist);
            blend_into_px(pixel, smp, C, blend, maxv);
            blend_byte = (uint8_t)sat_cast(roundf(blend * 255.0f), 255u);
        }
    }
    *blend_out = blend_byte;
    *warning = has_warning;
    if (st) {
        for (int i = 0; i < 4; ++i) { st->median[i] = median[i]; st->iqr_inv[i] = iqr_inv[i]; }
        st->num_outliers = k;
        st->max_index = max_index;
        st->max_dist_sq = max_dist_sq;
        st->bg_index = bg_index;
        st->warning = (uint32_t)has_warning;
        st->pad = 0;
    }
}

/* OutlierProcessor::process per-pixel loop, src/chrono.rs:169-191. */
int or_outlier(const or_view* v, int32_t frame_offset, const uint32_t* sample_pos,
               uint32_t n_sample, const or_outlier_params* p, void* out_px, uint8_t* out_mask,
               uint64_t* warnings, or_pixel_stats* stats, uint32_t y0, uint32_t y1,
               int threads) {
    const uint32_t S = sample_pos ? n_sample : v->n;
    if (S == 0) return -1;
    if (!p->threshold.absolute && S < 3) return -4;
    if (v->n == 0 || v->channels < 1 || v->channels > 4) return -1;
    if (y1 > v->out_h) y1 = v->out_h;
    const uint32_t C = v->channels, N = v->n, W = v->out_w;
    uint64_t warn_total = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(+ : warn_total)
#endif
    {
        scratch_t s;
        s.values = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (S + 1));
        s.sort_tmp = (uint32_t*)malloc(sizeof(uint32_t) * (S + 1));
        s.out_idx = (uint32_t*)malloc(sizeof(uint32_t) * N);
        s.out_dist = (float*)malloc(sizeof(float) * N);
        s.non_out = (uint32_t*)malloc(sizeof(uint32_t) * N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t yy = (int64_t)y0; yy < (int64_t)y1; ++yy) {
            for (uint32_t x = 0; x < W; ++x) {
                uint32_t pixel[4] = {0, 0, 0, 0};
                uint8_t blend;
                int warn;
                or_pixel_stats* st = stats ? &stats[(size_t)yy * W + x] : NULL;
                calc_pixel(v, x, (uint32_t)yy, frame_offset, sample_pos, S, p, &s, pixel,
                           &blend, &warn, st);
                if (warn) warn_total++;
                size_t o = ((size_t)yy * W + x) * C;
                for (uint32_t ch = 0; ch < C; ++ch) {
                    if (v->bytes_per_sample == 1) ((uint8_t*)out_px)[o + ch] = (uint8_t)pixel[ch];
                    else ((uint16_t*)out_px)[o + ch] = (uint16_t)pixel[ch];
                    if (out_mask) out_mask[o + ch] = ch < 3 ? blend : 255;
                }
            }
        }
        free(s.values); free(s.sort_tmp); free(s.out_idx); free(s.out_dist); free(s.non_out);
    }
    if (warnings) *warnings = warn_total;
    return 0;
}

/* ------------------------------------------------------------- simple.rs */

/* SimpleProcessor::process, src/simple.rs:26-168 (frames in index order, pixel loop
 * :96-133). */
int or_simple(const or_view* v, int32_t frame_offset, const float weights[4],
              const or_fade* fade, int32_t darker, void* out_px, uint32_t y0, uint32_t y1,
              int threads) {
    if (v->n == 0 || v->channels < 1 || v->channels > 4) return -1;
    if (y1 > v->out_h) y1 = v->out_h;
    const uint32_t C = v->channels, N = v->n, W = v->out_w;
    const uint32_t maxv = v->bytes_per_sample == 1 ? 255u : 65535u;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
#endif
    for (int64_t yy = (int64_t)y0; yy < (int64_t)y1; ++yy) {
        for (uint32_t x = 0; x < W; ++x) {
            uint32_t out[4] = {0, 0, 0, 0};
            float extreme = darker ? 3.40282347e+38f : -3.40282347e+38f;
            for (uint32_t t = 0; t < N; ++t) {
                uint32_t in[4];
                float value = 0.0f;
                for (uint32_t ch = 0; ch < C; ++ch) {
                    in[ch] = sample_at(v, t, x, (uint32_t)yy, ch);
                    value += (float)in[ch] * weights[ch];
                }
                int is_ext = darker ? (value < extreme) : (value > extreme);
                if (is_ext) {
                    extreme = value;
                    float f = or_fade_pick(fade, (int32_t)t, (int32_t)N, frame_offset);
                    if (f > 0.0f) {
                        if (f >= 1.0f) {
                            for (uint32_t ch = 0; ch < C; ++ch) out[ch] = in[ch];
                        } else {
                            blend_into_px(out, in, C, f, maxv);
                        }
                    }
                }
            }
            size_t o = ((size_t)yy * W + x) * C;
            for (uint32_t ch = 0; ch < C; ++ch) {
                if (v->bytes_per_sample == 1) ((uint8_t*)out_px)[o + ch] = (uint8_t)out[ch];
                else ((uint16_t*)out_px)[o + ch] = (uint16_t)out[ch];
            }
        }
    }
    return 0;
}

/* -------------------------------------------------------------- shake.rs */

/* ShakeAnalyzer::analyze + fill_windows + calc_diffs, src/shake.rs:190-385.  The view's
 * crop must be NULL (offsets are measured on uncropped frames). */
int or_shake(const or_view* v, const int32_t* anchors_xy, uint32_t n_anchors,
             uint32_t anchor_radius, uint32_t search_radius, int32_t* offsets, int threads) {
    const int32_t size = (int32_t)(2 * anchor_radius + 1);
    const int32_t ssize = (int32_t)(2 * search_radius + 1);
    const int32_t r = (int32_t)anchor_radius, sr = (int32_t)search_radius;
    const uint32_t C = v->channels;
    const int64_t W = v->width, H = v->height;
    /* bounds: fill_windows (:333-336) and calc_diffs (:371-374) panic out of image */
    for (uint32_t a = 0; a < n_anchors; ++a) {
        int64_t cx = anchors_xy[2 * a], cy = anchors_xy[2 * a + 1];
        if (cx - r < 0 || cy - r < 0 || cx + r >= W || cy + r >= H) return -3;
        if (v->n > 1 && (cx - r - sr < 0 || cy - r - sr < 0 || cx + r + sr >= W ||
                         cy + r + sr >= H))
            return -3;
    }
    const size_t win_len = (size_t)size * size * C;
    uint32_t* wins = (uint32_t*)malloc(sizeof(uint32_t) * win_len * (n_anchors ? n_anchors : 1));
    for (uint32_t a = 0; a < n_anchors; ++a) {
        int32_t cx = anchors_xy[2 * a], cy = anchors_xy[2 * a + 1];
        uint32_t* win = wins + a * win_len;
        for (int32_t dy = 0; dy < size; ++dy) {
            int32_t yy = cy + dy - r;
            for (int32_t dx = 0; dx < size; ++dx) {
                int32_t xx = cx + dx - r;
                size_t idx = ((size_t)dy * size + dx) * C;
                for (uint32_t ch = 0; ch < C; ++ch)
                    win[idx + ch] = sample_at(v, 0, (uint32_t)xx, (uint32_t)yy, ch);
            }
        }
    }
    offsets[0] = 0;
    offsets[1] = 0;
    const int wide = v->bytes_per_sample == 2;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
#endif
    for (int64_t f = 1; f < (int64_t)v->n; ++f) {
        int64_t* diff = (int64_t*)calloc((size_t)ssize * ssize, sizeof(int64_t));
        for (uint32_t a = 0; a < n_anchors; ++a) {
            int32_t cx = anchors_xy[2 * a], cy = anchors_xy[2 * a + 1];
            const uint32_t* win = wins + a * win_len;
            for (int32_t oy = 0; oy < ssize; ++oy) {
                for (int32_t ox = 0; ox < ssize; ++ox) {
                    int32_t di = oy * ssize + ox;
                    int64_t acc = diff[di];
                    for (int32_t dy = 0; dy < size; ++dy) {
                        int32_t yy = cy + (oy - sr) + dy - r;
                        for (int32_t dx = 0; dx < size; ++dx) {
                            int32_t xx = cx + (ox - sr) + dx - r;
                            size_t idx = ((size_t)dy * size + dx) * C;
                            for (uint32_t ch = 0; ch < C; ++ch) {
                                int64_t d = (int64_t)win[idx + ch] -
                                            (int64_t)sample_at(v, (uint32_t)f, (uint32_t)xx,
                                                               (uint32_t)yy, ch);
                                if (wide) acc += d * d;
                                else acc = (int64_t)(int32_t)((uint32_t)acc + (uint32_t)(d * d));
                            }
                        }
                    }
                    diff[di] = acc;
                }
            }
        }
        /* :294-297 first minimum in row-major (oy, ox) order */
        int32_t best = 0;
        for (int32_t i = 1; i < ssize * ssize; ++i)
            if (diff[i] < diff[best]) best = i;
        offsets[2 * f] = best % ssize - sr;
        offsets[2 * f + 1] = best / ssize - sr;
        free(diff);
    }
    free(wins);
    return 0;
}

/* Crop::create, src/shake.rs:136-176. */
int or_crop_create(const int32_t* offsets, uint32_t n, uint32_t width, uint32_t height,
                   int32_t* crop_xy, uint32_t* out_w, uint32_t* out_h) {
    int32_t xmin = 0, ymin = 0, xmax = 0, ymax = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t x = offsets[2 * i], y = offsets[2 * i + 1];
        if (x < xmin) xmin = x;
        if (y < ymin) ymin = y;
        if (x > xmax) xmax = x;
        if (y > ymax) ymax = y;
    }
    if (xmin == 0 && ymin == 0 && xmax == 0 && ymax == 0) {
        *out_w = width;
        *out_h = height;
        return 0;
    }
    *out_w = (uint32_t)((int32_t)width + xmin - xmax);
    *out_h = (uint32_t)((int32_t)height + ymin - ymax);
    for (uint32_t i = 0; i < n; ++i) {
        crop_xy[2 * i] = -xmin + offsets[2 * i];
        crop_xy[2 * i + 1] = -ymin + offsets[2 * i + 1];
    }
    return 1;
}

/* --------------------------------------------------------------- main.rs */

/* create_video window math, src/main.rs:230-286 (identical in create_video_simple
 * :349-404 with image_count = files.len()). */
int32_t or_video_windows(int32_t in_start_has, int32_t in_start, int32_t in_end_has,
                         int32_t in_end, uint32_t in_step, int32_t out_start_has,
                         int32_t out_start, int32_t out_end_has, int32_t out_end,
                         uint32_t out_step, uint32_t image_count, int32_t* win_start,
                         int32_t* win_end, int32_t* win_frame, int32_t cap) {
    int32_t v_lower;
    if (out_start_has) v_lower = out_start;
    else if (in_start_has && in_end_has) v_lower = -(in_end - in_start) + 1;
    else v_lower = 0;
    int32_t v_upper = out_end_has ? out_end : (int32_t)image_count;
    int32_t nframes = (v_upper - v_lower) / (int32_t)out_step;
    if (nframes < 0) nframes = 0;
    int32_t step = (int32_t)in_step;
    for (int32_t i = 0; i < nframes; ++i) {
        int32_t frame = i * (int32_t)out_step + v_lower;
        int32_t start, end;
        if (in_start_has) {
            int32_t st = frame + in_start;
            while (st < 0) st += step;
            int32_t a = st % step, b = frame + in_start;
            start = a > b ? a : b;
        } else {
            start = 0;
        }
        if (in_end_has) {
            int32_t a = (int32_t)image_count + (frame + in_end) % step - step;
            int32_t b = frame + in_end;
            end = a < b ? a : b;
        } else {
            end = (int32_t)image_count;
        }
        if (i < cap) {
            win_start[i] = start;
            win_end[i] = end;
            win_frame[i] = frame - v_lower;
        }
    }
    return nframes;
}
