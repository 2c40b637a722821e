/*
 * chrono_oracle.h -- CPU restatement of chrono-photo's per-pixel compositing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py may load it.  The product path (the HIP
 * library behind include/chrono_hip.h) never calls, links or falls back to it.
 *
 * Every function restates the reference crate chrono-photo v0.6.5 line by line
 * (citations are src/<file>.rs:<line> of that crate).  Arithmetic is IEEE f32 in the
 * reference's operation order; build with -ffp-contract=off and no fast-math.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   or_quantile / or_median       reference KAT src/chrono.rs:598-604, numpy weibull
 *                                 quantile (Hyndman-Fan type 6), brute force
 *   or_threshold_new/blend_value  closed form src/options.rs:197-231
 *   or_fade_table / or_fade_get   closed form (linear keyframes), clamp/repeat laws
 *   or_blend_into_u8              commented reference KAT src/color.rs:51-64
 *   or_outlier                    brute-force invariants (mask <=> dist^2 >= min^2,
 *                                 first/last/extreme selection, darker/lighter laws),
 *                                 numpy float32 re-evaluation of the distances; the
 *                                 background average with and without outliers against
 *                                 the exact rational of the reference's expression
 *                                 (blends switched off by a zero fade); the average,
 *                                 forward and backward modes at full blend against the
 *                                 exact rounded outlier mean and the last / first
 *                                 outlier; partial blends with per-frame fades of
 *                                 1/4, 1/2, 3/4 (exact in f32) against exact rational
 *                                 convex steps in frame / reverse order, and the
 *                                 average mode's frame-0 fade.  The f32 rounding of
 *                                 other partial blends follows the reference's
 *                                 a + (b - a) t and is checked by the blend KAT only.
 *   or_simple                     numpy argmin/argmax (first index on ties) without fade
 *   or_shake                      brute-force int64 SSD, recovery of seeded translations
 *   or_crop_create                closed form src/shake.rs:136-176
 *   or_video_windows              hand-derived window tables src/main.rs:230-286
 *   random background             values parity unpinned: the reference draws from an
 *                                 OS-seeded ThreadRng; this restatement uses a
 *                                 seeded counter hash.  Pinned: the candidate logic
 *                                 (support of the draws with and without outliers, the
 *                                 swap quirk) and uniform frequencies.
 */
#ifndef CHRONO_ORACLE_H
#define CHRONO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t absolute;
    float min, max, scale;
} or_threshold;

typedef struct {
    int32_t is_none;   /* Fade::none() */
    int32_t repeat;    /* FadeMode::Repeat (else Clamp) */
    int32_t absolute;
    int32_t offset;
    uint32_t len;      /* number of table entries */
    const float* values;
} or_fade;

enum { OR_BG_FIRST = 0, OR_BG_RANDOM = 1, OR_BG_AVERAGE = 2, OR_BG_MEDIAN = 3 };
enum { OR_SEL_FIRST = 0, OR_SEL_LAST = 1, OR_SEL_EXTREME = 2, OR_SEL_AVERAGE = 3,
       OR_SEL_FORWARD = 4, OR_SEL_BACKWARD = 5 };

typedef struct {
    or_threshold threshold;
    int32_t background;
    int32_t outlier;
    float weights[4];
    or_fade fade;
    uint64_t seed;      /* random background only */
} or_outlier_params;

/* Per-pixel diagnostics (optional output of or_outlier). */
typedef struct {
    float median[4];
    float iqr_inv[4];
    uint32_t num_outliers;
    uint32_t max_index;
    float max_dist_sq;
    uint32_t bg_index;      /* first/random background: chosen frame position */
    uint32_t warning;
    uint32_t pad;
} or_pixel_stats;

/* Frame view: a contiguous stack [slots][height][width][channels] of 1- or 2-byte
 * samples; `slots[i]` picks the frame at position i; `crop_xy` (nullable) gives the
 * per-position crop origin; the output is out_w x out_h. */
typedef struct {
    const void* data;
    uint32_t bytes_per_sample;
    uint32_t channels;
    uint32_t width, height;
    const uint32_t* slots;    /* nullable: identity */
    uint32_t n;               /* frames in this composite */
    const int32_t* crop_xy;   /* nullable */
    uint32_t out_w, out_h;
} or_view;

void or_threshold_new(int32_t absolute, float min, float max, float range, or_threshold* out);
float or_blend_value(const or_threshold* t, float dist);

/* Fade::new table from (frame, value) keyframes; returns table length (last-first+1),
 * writes at most cap values. */
uint32_t or_fade_table(const int32_t* frames, const float* values, uint32_t nkeys,
                       float* out, uint32_t cap);
float or_fade_get(const or_fade* f, int32_t frame);
float or_fade_pick(const or_fade* f, int32_t frame, int32_t total, int32_t offset);

float or_median(const uint32_t* sorted, uint32_t n);
float or_quantile(const uint32_t* sorted, uint32_t n, float q);
void or_blend_into_u8(uint8_t* a, const uint8_t* b, uint32_t n, float blend);

/* Outlier composite over rows [y0, y1) of the output.  sample_pos (nullable) lists the
 * sorted positions used for the median/IQR.  Returns 0, or -4 when rel thresholds meet
 * fewer than 3 samples (the reference panics there). */
int or_outlier(const or_view* v, int32_t frame_offset, const uint32_t* sample_pos,
               uint32_t n_sample, const or_outlier_params* p, void* out_px,
               uint8_t* out_mask, uint64_t* warnings, or_pixel_stats* stats,
               uint32_t y0, uint32_t y1, int threads);

/* Darker/lighter composite (src/simple.rs:26-176) over rows [y0, y1). */
int or_simple(const or_view* v, int32_t frame_offset, const float weights[4],
              const or_fade* fade, int32_t darker, void* out_px, uint32_t y0, uint32_t y1,
              int threads);

/* Shake offsets (src/shake.rs:190-385): offsets[2*i] = dx, [2*i+1] = dy for each
 * position i (position 0 is the reference frame).  Returns -3 on out-of-image
 * coordinates (the reference panics). */
int or_shake(const or_view* v, const int32_t* anchors_xy, uint32_t n_anchors,
             uint32_t anchor_radius, uint32_t search_radius, int32_t* offsets, int threads);

/* Crop::create (src/shake.rs:136-176).  Returns 1 and fills crops when a crop is needed,
 * 0 when every offset is zero (the reference returns None). */
int or_crop_create(const int32_t* offsets, uint32_t n, uint32_t width, uint32_t height,
                   int32_t* crop_xy, uint32_t* out_w, uint32_t* out_h);

/* Video window table (src/main.rs:230-286).  *_has = 0 means "." (None).
 * Writes per output frame: start index into win_start, count into win_count and frame
 * number (frame - v_lower) into win_frame; indices follow start, start+step, ...
 * Returns the number of output frames (windows, including empty ones). */
int32_t or_video_windows(int32_t in_start_has, int32_t in_start, int32_t in_end_has,
                         int32_t in_end, uint32_t in_step, int32_t out_start_has,
                         int32_t out_start, int32_t out_end_has, int32_t out_end,
                         uint32_t out_step, uint32_t image_count, int32_t* win_start,
                         int32_t* win_end, int32_t* win_frame, int32_t cap);

/* Counter-hash RNG used for the random background (seed, pixel, draw) -> [0, n). */
uint32_t or_rng_range(uint64_t seed, uint64_t pixel, uint32_t draw, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
