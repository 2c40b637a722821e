// synth.h -- seeded synthetic frame stacks for tests and benchmarks.
//
// Integer-only arithmetic, so the host and the device produce identical bytes.
// A scene is a horizontal gradient plus approximately Gaussian sensor noise (sum of four
// uniform 16-bit draws, Irwin-Hall), optional moving dark disks, optional bright blobs on
// a fraction of frames, an optional static block texture in world coordinates and an
// optional per-frame integer translation (camera shake).  Frame t shows the world shifted
// by (dx_t, dy_t): frame_t(x, y) = world(x - dx_t, y - dy_t) + noise_t(x, y).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CS_HD __host__ __device__ inline
#else
#define CS_HD static inline
#endif

typedef struct {
    uint64_t seed;
    uint32_t width, height, channels, bytes_per_sample, n_frames;
    int32_t base_lo, base_hi;      // gradient endpoints, sample units
    int32_t sigma_x16;             // noise sigma * 16
    uint32_t n_disks;              // moving dark disks
    int32_t disk_r_min, disk_r_max;
    int32_t disk_val_lo, disk_val_hi;
    uint32_t blob_permille;        // frames carrying a bright blob, per mille
    int32_t blob_r, blob_val;
    int32_t shake_max;             // per-frame translation in [-shake_max, shake_max]
    int32_t texture_amp;           // static world-space 8x8 block texture amplitude
    int32_t disk_y_lo, disk_y_hi;  // row band of the disk paths (0, 0: the whole frame)
} cs_scene;

CS_HD uint64_t cs_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

CS_HD uint64_t cs_hash(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return cs_mix(seed ^ cs_mix(a * 0x100000001B3ull ^ cs_mix(b * 0xC2B2AE3D27D4EB4Full ^ c)));
}

// uniform integer in [lo, hi]
CS_HD int32_t cs_uniform(uint64_t h, int32_t lo, int32_t hi) {
    uint64_t span = (uint64_t)(int64_t)(hi - lo + 1);
    return lo + (int32_t)(((h >> 32) * span) >> 32);
}

CS_HD void cs_shake_offset(const cs_scene* s, uint32_t t, int32_t* dx, int32_t* dy) {
    if (t == 0 || s->shake_max <= 0) { *dx = 0; *dy = 0; return; }
    uint64_t h = cs_hash(s->seed, 0x5A4Bull, t, 0);
    *dx = cs_uniform(h, -s->shake_max, s->shake_max);
    *dy = cs_uniform(cs_mix(h), -s->shake_max, s->shake_max);
}

// Disk d: radius and a linear path from (x0,y0) at frame 0 to (x1,y1) at frame N-1.
CS_HD void cs_disk(const cs_scene* s, uint32_t d, int32_t* r, int32_t* x0, int32_t* y0,
                   int32_t* x1, int32_t* y1) {
    uint64_t h = cs_hash(s->seed, 0xD15Cull, d, 0);
    *r = cs_uniform(h, s->disk_r_min, s->disk_r_max);
    int32_t W = (int32_t)s->width, H = (int32_t)s->height;
    const int32_t ylo = s->disk_y_hi > s->disk_y_lo ? s->disk_y_lo : 0;
    const int32_t yhi = s->disk_y_hi > s->disk_y_lo ? s->disk_y_hi : H - 1;
    h = cs_mix(h); *y0 = cs_uniform(h, ylo + *r, yhi - *r);
    h = cs_mix(h); *y1 = cs_uniform(h, ylo + *r, yhi - *r);
    *x0 = *r + W / 16;
    *x1 = W - 1 - *r - W / 16;
    if (d & 1u) { int32_t t = *x0; *x0 = *x1; *x1 = t; }
}

CS_HD int32_t cs_noise(const cs_scene* s, uint32_t t, uint32_t x, uint32_t y, uint32_t ch) {
    uint64_t h = cs_hash(s->seed, ((uint64_t)t << 2) | ch, x, y);
    int64_t sum = (int64_t)(h & 0xFFFF) + (int64_t)((h >> 16) & 0xFFFF) +
                  (int64_t)((h >> 32) & 0xFFFF) + (int64_t)(h >> 48);
    int64_t n = sum - 131070;  // zero mean, sd 37837.6
    // noise * 256 = n / 37838 * sigma ; sigma = sigma_x16 / 16
    return (int32_t)((n * (int64_t)s->sigma_x16 * 16) / 37838);
}

// One sample of frame t at (x, y), channel ch.
CS_HD uint32_t cs_sample(const cs_scene* s, uint32_t t, uint32_t x, uint32_t y, uint32_t ch) {
    const int32_t maxv = s->bytes_per_sample == 1 ? 255 : 65535;
    int32_t dx, dy;
    cs_shake_offset(s, t, &dx, &dy);
    const int32_t wx = (int32_t)x - dx, wy = (int32_t)y - dy;  // world coordinates
    const int32_t W = (int32_t)s->width;
    int64_t v256 = (int64_t)s->base_lo * 256 +
                   ((int64_t)(s->base_hi - s->base_lo) * 256 * wx) / (W > 1 ? W - 1 : 1);
    if (s->texture_amp > 0) {
        uint64_t hb = cs_hash(s->seed, 0x7E47ull, (uint64_t)(int64_t)(wx >> 3),
                              (uint64_t)(int64_t)(wy >> 3));
        v256 += (int64_t)cs_uniform(hb, -s->texture_amp, s->texture_amp) * 256;
    }
    const int32_t N = (int32_t)s->n_frames;
    for (uint32_t d = 0; d < s->n_disks; ++d) {
        int32_t r, x0, y0, x1, y1;
        cs_disk(s, d, &r, &x0, &y0, &x1, &y1);
        int32_t den = N > 1 ? N - 1 : 1;
        int32_t cx = x0 + (int32_t)(((int64_t)(x1 - x0) * (int32_t)t) / den);
        int32_t cy = y0 + (int32_t)(((int64_t)(y1 - y0) * (int32_t)t) / den);
        int64_t ex = wx - cx, ey = wy - cy;
        if (ex * ex + ey * ey <= (int64_t)r * r) {
            uint64_t hv = cs_hash(s->seed, 0xDA7Aull, d, t);
            v256 = (int64_t)cs_uniform(hv, s->disk_val_lo, s->disk_val_hi) * 256;
        }
    }
    if (s->blob_permille > 0) {
        uint64_t hb = cs_hash(s->seed, 0xB10Bull, t, 0);
        if ((uint32_t)cs_uniform(hb, 0, 999) < s->blob_permille) {
            int32_t H = (int32_t)s->height;
            int32_t bx = cs_uniform(cs_mix(hb), s->blob_r, W - 1 - s->blob_r);
            int32_t by = cs_uniform(cs_mix(cs_mix(hb)), s->blob_r, H - 1 - s->blob_r);
            int64_t ex = wx - bx, ey = wy - by;
            if (ex * ex + ey * ey <= (int64_t)s->blob_r * s->blob_r)
                v256 = (int64_t)s->blob_val * 256;
        }
    }
    v256 += cs_noise(s, t, x, y, ch);
    int64_t v = (v256 + 128) >> 8;
    if (v < 0) v = 0;
    if (v > maxv) v = maxv;
    return (uint32_t)v;
}
