// kparams.h -- kernel argument blocks (plain data, shared by host and device code).
#pragma once
#include <stdint.h>

namespace cp {

struct ThresholdK {
    int32_t absolute;
    float min, max, scale;
};

struct FadeK {
    int32_t is_none, repeat, absolute, offset;
    uint32_t len;
    const float* values;  // device table
};

// Per-pixel diagnostics; layout shared with the C ABI (cp_pixel_stats).
struct PixelStats {
    float median[4];
    float iqr_inv[4];
    uint32_t num_outliers;
    uint32_t max_index;
    float max_dist_sq;
    uint32_t bg_index;
    uint32_t warning;
    uint32_t pad;
};

// One composite over a device-resident stack.  Frame position t (0..n-1) reads sample
// (x, y, ch) at element index pos_off[t] + (y * stack_w + x) * C + ch, where pos_off folds
// the stack slot and the per-frame crop origin.
struct OutlierK {
    const uint8_t* stack;
    const uint64_t* pos_off;     // n element offsets
    const uint32_t* sample_mask; // bit t set: position t is in the median sample (null: all)
    const uint64_t* samp_off;    // element offsets of the n_sample sample positions (null: all)
    uint32_t n, n_sample;
    uint32_t stack_w;            // row pitch in pixels
    uint32_t out_w, out_h;
    int32_t frame_offset;
    ThresholdK th;
    int32_t background, outlier;
    float w[4];
    FadeK fade;
    uint64_t seed;
    uint64_t pix_base;           // image index of output pixel (0, 0): row bands draw the same random numbers
    void* out_px;
    uint8_t* out_mask;
    unsigned long long* warnings;
    PixelStats* stats;
};

struct SimpleK {
    const uint8_t* stack;
    const uint64_t* pos_off;
    uint32_t n;
    uint32_t stack_w;
    uint32_t out_w, out_h;
    int32_t frame_offset;
    int32_t darker;
    float w[4];
    FadeK fade;
    void* out_px;
};

}  // namespace cp
