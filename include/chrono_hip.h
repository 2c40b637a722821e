/*
 * chrono_hip.h -- C ABI of the MI355X-native chrono-photo compositing path.
 *
 * Drop-in boundary for the per-pixel stack compositing of chrono-photo v0.6.5 (a Rust
 * crate).  Each entry point replaces a reference interface, cited as src/<file>.rs:<line>
 * of that crate.  The reference is a library + CLI that reads image files; on this side
 * the frames live in a device-resident stack (HBM) and the caller passes plain host or
 * device pointers.  No entry point throws or aborts: each returns a cp_status and
 * cp_last_error() gives a thread-local message.  A cp_ctx is not thread-safe; use one per
 * host thread, or the server API.
 */
#ifndef CHRONO_HIP_H
#define CHRONO_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CP_API __attribute__((visibility("default")))

/* Status codes.  The reference panics or returns io/image errors at these points. */
typedef enum {
    CP_OK = 0,
    CP_ERR_INVALID_ARG = -1,         /* bad enum, size, null pointer */
    CP_ERR_LAYOUT_MISMATCH = -2,     /* frame size differs: panic at src/simple.rs:62-66 */
    CP_ERR_ANCHOR_OUT_OF_RANGE = -3, /* panic at src/shake.rs:333-336, :371-374 */
    CP_ERR_TOO_FEW_FRAMES_REL = -4,  /* rel threshold, <3 samples: usize underflow src/chrono.rs:569 */
    CP_ERR_OUT_OF_HBM_BUDGET = -5,   /* allocation would exceed the context's HBM budget */
    CP_ERR_HIP = -6,                 /* HIP runtime error */
    CP_ERR_NO_DEVICE = -7,           /* no usable GPU */
    CP_ERR_CAPACITY = -8             /* stack or queue full */
} cp_status;

typedef struct cp_ctx cp_ctx;
typedef struct cp_stack cp_stack;
typedef struct cp_server cp_server;

/* Threshold, src/options.rs:189-231.  Fields hold Threshold::new's pre-scaled values. */
typedef struct {
    int32_t absolute;
    float min, max, scale;
} cp_threshold;

/* Fade, src/options.rs:60-140.  values: table built by cp_fade_table (host memory). */
typedef struct {
    int32_t is_none;  /* Fade::none() */
    int32_t repeat;   /* FadeMode::Repeat, else Clamp */
    int32_t absolute;
    int32_t offset;   /* first keyframe */
    uint32_t len;
    const float* values;
} cp_fade;

/* BackgroundMode, src/options.rs:319-345. */
enum { CP_BG_FIRST = 0, CP_BG_RANDOM = 1, CP_BG_AVERAGE = 2, CP_BG_MEDIAN = 3 };
/* OutlierSelectionMode, src/options.rs:284-316. */
enum { CP_SEL_FIRST = 0, CP_SEL_LAST = 1, CP_SEL_EXTREME = 2, CP_SEL_AVERAGE = 3,
       CP_SEL_FORWARD = 4, CP_SEL_BACKWARD = 5 };
/* SelectionMode, src/options.rs:8-32. */
enum { CP_MODE_OUTLIER = 0, CP_MODE_LIGHTER = 1, CP_MODE_DARKER = 2 };

/* OutlierProcessor::new arguments, src/chrono.rs:46-54.  sample_pos replaces the
 * reference's ThreadRng draw of `--sample` indices (src/chrono.rs:151-163): the caller
 * passes the sorted positions (NULL: all frames).  seed drives the random background. */
typedef struct {
    cp_threshold threshold;
    int32_t background;
    int32_t outlier;
    float weights[4];
    cp_fade fade;
    uint64_t seed;
    const uint32_t* sample_pos;
    uint32_t n_sample;
} cp_outlier_params;

/* Frames of one composite: the reference's `image_indices` (src/chrono.rs:79,101) plus the
 * per-frame crops of shake correction (src/shake.rs:136-180).  frames == NULL means slots
 * 0..n-1.  crop_xy (nullable) holds the crop origin (x, y) of each position; out_width and
 * out_height give the cropped size (0: the stack size). */
typedef struct {
    const uint32_t* frames;
    uint32_t n;
    int32_t frame_offset;  /* fade offset: image_indices[0] in the reference */
    const int32_t* crop_xy;
    uint32_t out_width, out_height;
} cp_frames;

/* Per-pixel diagnostics of an outlier composite (medians, inverse IQR, outlier count,
 * extreme frame, background frame). */
typedef struct {
    float median[4];
    float iqr_inv[4];
    uint32_t num_outliers;
    uint32_t max_index;
    float max_dist_sq;
    uint32_t bg_index;
    uint32_t warning;
    uint32_t pad;
} cp_pixel_stats;

/* Output buffers.  With CP_FLAG_DEVICE_OUT all pointers are device pointers and the call
 * is asynchronous on the context stream; warnings then points to a device u64 that the
 * kernel adds to.  Otherwise they are host pointers and the call returns when done. */
typedef struct {
    void* px;               /* out_h * out_w * C samples (u8 or u16) */
    uint8_t* mask;          /* nullable: blend mask, src/chrono.rs:183-191 */
    uint64_t* warnings;     /* nullable: all-outlier pixel count, src/chrono.rs:180-203 */
    cp_pixel_stats* stats;  /* nullable */
} cp_outputs;

#define CP_FLAG_DEVICE_OUT 1u
#define CP_FLAG_FORCE_GENERIC 2u /* use the one-lane-per-pixel kernel (validation) */

typedef struct {
    uint32_t width, height, channels, bytes_per_sample;
    uint32_t capacity, count;
    void* device_ptr;
} cp_stack_info;

/* ------------------------------------------------------------------ runtime */
CP_API const char* cp_last_error(void);
CP_API const char* cp_version(void);
CP_API int cp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, cp_ctx** out);
CP_API int cp_ctx_destroy(cp_ctx* ctx);
CP_API int cp_ctx_synchronize(cp_ctx* ctx);
CP_API void* cp_ctx_stream(cp_ctx* ctx);

/* Device-resident frame stack [capacity][height][width][channels] replacing the time
 * slicer's gzip slice files (src/slicer.rs:106-233, src/streams.rs:87-220). */
CP_API int cp_stack_create(cp_ctx* ctx, uint32_t width, uint32_t height, uint32_t channels,
                           uint32_t bytes_per_sample, uint32_t capacity, cp_stack** out);
/* Borrow caller-owned device memory holding n_frames contiguous frames. */
CP_API int cp_stack_wrap(cp_ctx* ctx, void* device_ptr, uint32_t width, uint32_t height,
                         uint32_t channels, uint32_t bytes_per_sample, uint32_t n_frames,
                         cp_stack** out);
/* Copy n host frames into slots [slot, slot+n); count grows to cover them. */
CP_API int cp_stack_upload(cp_stack* st, uint32_t slot, const void* host_frames, uint32_t n);
/* Append n host frames; returns the first slot used (layout check as src/slicer.rs:145-157). */
CP_API int cp_stack_push(cp_stack* st, const void* host_frames, uint32_t n, uint32_t* first_slot);
/* Declare slots [0, count) filled by a device-side writer (a decoder or generator writing
 * through cp_stack_info.device_ptr on the context stream). */
CP_API int cp_stack_commit(cp_stack* st, uint32_t count);
/* Copy frames [slot, slot+n) back to host memory (the inverse of cp_stack_upload). */
CP_API int cp_stack_download(const cp_stack* st, uint32_t slot, void* host_frames, uint32_t n);
CP_API int cp_stack_info_get(const cp_stack* st, cp_stack_info* out);
CP_API int cp_stack_destroy(cp_stack* st);

/* ------------------------------------------------------------ option values */
/* Threshold::new, src/options.rs:197-214 (range 255 for u8 samples, 65535 for u16). */
CP_API int cp_threshold_new(int32_t absolute, float min, float max, uint32_t bytes_per_sample,
                            cp_threshold* out);
/* Threshold::blend_value, src/options.rs:224-231. */
CP_API float cp_threshold_blend_value(const cp_threshold* t, float dist);
/* Fade::new table, src/options.rs:69-95; *len receives the table length. */
CP_API int cp_fade_table(const int32_t* frames, const float* values, uint32_t nkeys, float* out,
                         uint32_t cap, uint32_t* len);
/* Fade::get, src/options.rs:113-140. */
CP_API float cp_fade_get(const cp_fade* fade, int32_t frame);

/* --------------------------------------------------------------- compositing */
/* OutlierProcessor::process per-pixel path (src/chrono.rs:73-206 + calc_pixel :208-494). */
CP_API int cp_outlier(cp_ctx* ctx, const cp_stack* st, const cp_frames* frames,
                      const cp_outlier_params* params, const cp_outputs* out, uint32_t flags);
/* Out-of-core OutlierProcessor::process: the reference composites time slices that its
 * slicer spilled to disk (src/chrono.rs:96-205 over src/slicer.rs, src/streams.rs:115-219);
 * here the frames stay in host memory and row bands of every frame position are streamed
 * through a device band stack, one composite per band.  host_frames[i] points at frame i
 * ([height][width][channels] samples, bytes_per_sample each); sel->frames (NULL: 0..n-1)
 * indexes host_frames, crops and frame_offset as for cp_outlier.  The band stack uses at
 * most band_bytes of device memory (0: what the context's HBM budget leaves, or a quarter
 * of the free device memory when the context has no budget).  Outputs are host buffers.
 * Results equal cp_outlier over a stack that holds the same frames. */
CP_API int cp_outlier_host(cp_ctx* ctx, const void* const* host_frames, uint32_t n_host_frames,
                           uint32_t width, uint32_t height, uint32_t channels,
                           uint32_t bytes_per_sample, const cp_frames* sel,
                           const cp_outlier_params* params, const cp_outputs* out, uint64_t band_bytes);
/* SimpleProcessor::process (src/simple.rs:26-168): darker != 0 selects `--mode darker`. */
CP_API int cp_simple(cp_ctx* ctx, const cp_stack* st, const cp_frames* frames,
                     const float weights[4], const cp_fade* fade, int32_t darker, void* out_px,
                     uint32_t flags);
/* SimpleProcessor::process over frames in host memory, streamed in row bands as for
 * cp_outlier_host; out_px is a host buffer.  Results equal cp_simple. */
CP_API int cp_simple_host(cp_ctx* ctx, const void* const* host_frames, uint32_t n_host_frames,
                          uint32_t width, uint32_t height, uint32_t channels,
                          uint32_t bytes_per_sample, const cp_frames* sel, const float weights[4],
                          const cp_fade* fade, int32_t darker, void* out_px, uint64_t band_bytes);
/* ShakeAnalyzer::analyze (src/shake.rs:190-313) over frame positions; offsets receives
 * (dx, dy) per position, position 0 = (0, 0).  frames may be NULL (slots 0..n-1).
 * Anchors whose window or search region leaves the image fail with
 * CP_ERR_ANCHOR_OUT_OF_RANGE (the reference panics); CP_ERR_INVALID_ARG when one window
 * row against one candidate row (about 12r + 12s bytes of region row, 12r of window row and
 * 8(2s+1) of box sums for 3 channels) exceeds 64 KiB of LDS: r up to about 1,800 or s up to
 * about 2,300 for RGB. */
CP_API int cp_shake(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                    const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                    uint32_t search_radius, int32_t* offsets);
/* cp_shake without any device allocation (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same search and results as cp_shake; the per-frame
 * candidate sums live in one static 8 MiB device array per GPU, so (8 n ((2s+1)^2) + 8,
 * rounded up to 256, plus 8 n for host outputs) must fit 8 MiB, else CP_ERR_CAPACITY
 * (s = 12: up to about 1,600 frames).  With CP_FLAG_DEVICE_OUT, offsets is a device array of
 * 2 n int32 written asynchronously on the context stream; otherwise a host array and the call
 * returns when done.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (other contexts or threads) race and must be serialised by the caller. */
CP_API int cp_shake_noalloc(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                            const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                            uint32_t search_radius, int32_t* offsets, uint32_t flags);
/* Crop::create (src/shake.rs:136-176): *has_crop = 0 where the reference returns None. */
CP_API int cp_crop_create(const int32_t* offsets, uint32_t n, uint32_t width, uint32_t height,
                          int32_t* crop_xy, uint32_t* out_w, uint32_t* out_h, int32_t* has_crop);

/* FrameRange (src/flist.rs:11-45): has_* = 0 stands for "." (None). */
typedef struct {
    int32_t has_start, start, has_end, end;
    uint32_t step;
} cp_frame_range;

/* create_video / create_video_simple window math (src/main.rs:230-286, :349-404).
 * Writes [start, end) and the output frame number of each window; *n_out = windows. */
CP_API int cp_video_windows(const cp_frame_range* video_in, const cp_frame_range* video_out,
                            uint32_t image_count, int32_t* win_start, int32_t* win_end,
                            int32_t* win_frame, int32_t cap, int32_t* n_out);

/* ------------------------------------------------------------------- batching
 * Many composites in a few launches (no reference counterpart: the reference runs one
 * processor per output, src/main.rs:260-330).  Outputs are device pointers. */
typedef struct {
    int32_t mode;               /* CP_MODE_* */
    const cp_stack* stack;
    cp_frames frames;
    cp_outlier_params outlier;  /* CP_MODE_OUTLIER */
    float weights[4];           /* darker/lighter */
    cp_fade fade;               /* darker/lighter */
    cp_outputs out;             /* device pointers */
} cp_job;

typedef struct cp_batch cp_batch;

/* Validate a list of jobs, stage all their per-call tables in device memory owned by the
 * batch, and capture the launch sequence in a HIP graph.  Each outlier job is planned as its
 * own cp_outlier call would run it (the context's tuned kernel and tile shape for its sample
 * width, channels, frame count and threshold kind).  Jobs planned on the same register-
 * resident kernel shape -- 8-bit stacks with 1-4 channels up to 1280 frames, with or without a
 * sample subset, 16-bit 3-4 channel stacks up to 512 frames -- share one launch whose grid
 * covers all their tiles, biggest jobs first.  Jobs left to the one-lane-per-pixel kernel and
 * darker/lighter jobs are launched one by one.  The launches are spread, biggest first, over
 * the context stream and three side streams, so the graph's branches run side by side.  The
 * caller's arrays are copied; stacks and output pointers must stay valid while the batch
 * exists. */
CP_API int cp_batch_create(cp_ctx* ctx, const cp_job* jobs, uint32_t n_jobs, uint32_t flags,
                           cp_batch** out);
/* Replay the batch on the context stream (asynchronous). */
CP_API int cp_batch_launch(cp_batch* batch);
CP_API int cp_batch_destroy(cp_batch* batch);
/* How a batch launches its jobs: group launches (one per kernel shape), the jobs they cover,
 * and the jobs launched one by one. */
typedef struct {
    uint32_t group_launches;
    uint32_t group_jobs;
    uint32_t direct_jobs;
    uint32_t reserved;
} cp_batch_info;
CP_API int cp_batch_info_get(const cp_batch* batch, cp_batch_info* out);
/* create + launch + wait + destroy. */
CP_API int cp_batch_run(cp_ctx* ctx, const cp_job* jobs, uint32_t n_jobs, uint32_t flags);
#define CP_FLAG_NO_GRAPH 4u /* cp_batch_create: launch kernels directly, no graph capture */

/* Autotuner: for each threshold kind (absolute, relative) and frame-count bucket of the
 * register-resident outlier kernels (<= 16, 32, 64, 128, 256, 512, 768, 1024, 1280 frames:
 * CP_TILE_BUCKETS buckets) time the candidate kernel choices on a synthetic stack of this
 * sample width and channel count (1-4) and keep the fastest for the context.  A choice is a tile shape of the 16-bit-halves kernel or of
 * the 8-bit packed-byte kernel, whose one-lane-per-pixel shapes rank either by the byte
 * search or by a sorting network; shapes (nullable, 2 * CP_TILE_BUCKETS entries: absolute
 * thresholds' buckets, then relative thresholds') receives
 * (kernel << 24) | (pixels_per_lane << 16) | (lanes_per_pixel << 8) | frames_per_lane per
 * bucket (kernel 0: 16-bit halves, 1: packed bytes with the byte search, 2: packed bytes
 * with the sorting network), or 0 where no register-resident kernel covers the bucket. */
#define CP_TILE_BUCKETS 9
CP_API int cp_autotune(cp_ctx* ctx, uint32_t bytes_per_sample, uint32_t channels, int32_t* shapes);
/* Sets the context's kernel choice per bucket without timing anything (no reference
 * counterpart): shapes holds 2 * CP_TILE_BUCKETS entries in cp_autotune's encoding, for instance
 * a result that cp_autotune returned in an earlier process, or 0 for the built-in default of
 * that bucket.  Each entry must be one of the bucket's candidates (CP_ERR_INVALID_ARG
 * otherwise, and nothing is changed). */
CP_API int cp_tile_shapes_set(cp_ctx* ctx, uint32_t bytes_per_sample, uint32_t channels, const int32_t* shapes);

/* -------------------------------------------------------------------- serving
 * Continuous-batching request loop (no reference counterpart: the reference's video mode
 * re-reads and recomputes every window from disk, src/main.rs:254-330).  A worker thread
 * admits queued requests in arrival order while the server's HBM budget and its
 * max_inflight stream slots allow, runs each on its own stream (upload, composite,
 * download), so admitted requests overlap on the device, and retires them as their
 * completion events fire.  Requests on resident stacks (e.g. the windows of a chrono-video
 * over one uploaded frame sequence) need no upload. */
typedef struct {
    cp_job job;               /* job.stack NULL: the request brings host frames below */
    const void* host_frames;  /* [n_frames][height][width][channels] */
    uint32_t width, height, channels, bytes_per_sample, n_frames;
    void* host_out_px;        /* nullable: copy the image to host memory when done */
    uint8_t* host_out_mask;   /* nullable */
} cp_request;

typedef struct {
    uint64_t submitted, completed, rejected;
    uint64_t pixel_samples;   /* frames * out_h * out_w over completed requests */
    uint64_t peak_bytes;      /* highest admitted HBM use */
    uint32_t peak_inflight;
    uint32_t kernel_downloads; /* results written to pinned host memory by a copy kernel
                                  (while uploads were in flight) instead of the DMA path */
} cp_server_stats;

CP_API int cp_server_create(cp_ctx* ctx, uint64_t hbm_budget_bytes, uint32_t max_inflight,
                            cp_server** out);
/* Queue a request; the caller's arrays must stay valid until the request completes. */
CP_API int cp_server_submit(cp_server* srv, const cp_request* req, uint64_t* ticket);
CP_API int cp_server_poll(cp_server* srv, uint64_t ticket, int32_t* done);
/* Block until the request is done; latency_ms = submit to completion.  Returns the
 * request's own status. */
CP_API int cp_server_wait(cp_server* srv, uint64_t ticket, double* latency_ms);
CP_API int cp_server_stats_get(cp_server* srv, cp_server_stats* out);
/* Drains in-flight work, then stops the worker. */
CP_API int cp_server_destroy(cp_server* srv);

#ifdef __cplusplus
}
#endif
#endif
