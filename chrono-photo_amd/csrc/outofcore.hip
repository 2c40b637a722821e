// outofcore.hip -- cp_outlier_host / cp_simple_host: composites of frame stacks that stay in
// host memory, streamed through the device in row bands.
//
// The reference never holds a whole stack in memory: its slicer writes time slices to disk
// and OutlierProcessor::process composites them one after another (src/chrono.rs:96-205,
// src/slicer.rs, src/streams.rs:115-219); SimpleProcessor streams whole frames
// (src/simple.rs:40-150).  Here a band of output rows [r0, r0 + rows) needs, from every
// frame position t, input rows [r0 + cy_t, r0 + rows + cy_t) (cy_t: the position's crop
// origin); the band stack holds rows [r0 + min cy, r0 + rows + max cy) of each position, so
// one frame-offset table serves every band and each band is composited by the same kernels
// as a resident stack.  The random background draws by image pixel index
// (OutlierK::pix_base), so a band produces exactly the pixels the whole-image composite would.
#include <algorithm>
#include <vector>

#include "runtime.h"

using namespace cp;

namespace {

// A band stack in device memory and the frame-offset table shared by all its bands.
struct BandPlan {
    std::vector<uint32_t> slot;  // host frame of each position
    int64_t min_cy = 0, halo = 0;
    uint64_t row_bytes = 0;      // one row of one frame
    uint32_t bh = 0;             // output rows per band
    cp_stack band;               // band.height = bh + halo rows per position
    std::vector<uint64_t> pos;   // element offset of each position's first output row
    uint64_t max_pos = 0;
    bool allocated = false;
};

// The host frames seen as a stack, for the geometry, frame and crop checks of the
// resident entry points.
cp_stack host_view(cp_ctx* ctx, uint32_t n, uint32_t W, uint32_t H, uint32_t C, uint32_t B) {
    cp_stack v;
    v.ctx = ctx;
    v.width = W;
    v.height = H;
    v.channels = C;
    v.bps = B;
    v.capacity = v.count = n;
    return v;
}

int check_host_args(cp_ctx* ctx, const void* const* host_frames, const cp_frames* sel, uint32_t n, uint32_t W,
                    uint32_t H, uint32_t C, uint32_t B) {
    if (!ctx || !host_frames || !sel) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (W == 0 || H == 0 || n == 0) return fail(CP_ERR_INVALID_ARG, "empty stack geometry");
    if (C < 1 || C > 4) return fail(CP_ERR_INVALID_ARG, "channels must be 1..4");
    if (B != 1 && B != 2) return fail(CP_ERR_INVALID_ARG, "bytes_per_sample must be 1 or 2");
    return CP_OK;
}

// Sizes and allocates the band stack for npos positions and out_h output rows from what
// band_bytes (0: the context's budget, or a quarter of the free device memory) leaves.
// Call with the context lock held and the outputs already allocated.
int plan_bands(cp_ctx* ctx, const void* const* host_frames, const cp_frames* sel, uint32_t W, uint32_t C,
               uint32_t B, uint32_t npos, uint32_t out_h, uint64_t band_bytes, BandPlan* bp) {
    bp->slot.resize(npos);
    std::vector<int64_t> cx(npos, 0), cy(npos, 0);
    for (uint32_t t = 0; t < npos; ++t) {
        bp->slot[t] = sel->frames ? sel->frames[t] : t;
        if (!host_frames[bp->slot[t]]) return fail(CP_ERR_INVALID_ARG, "host frame %u is null", bp->slot[t]);
        if (sel->crop_xy) {
            cx[t] = sel->crop_xy[2 * t];
            cy[t] = sel->crop_xy[2 * t + 1];
        }
    }
    bp->min_cy = *std::min_element(cy.begin(), cy.end());
    bp->halo = *std::max_element(cy.begin(), cy.end()) - bp->min_cy;
    bp->row_bytes = (uint64_t)W * C * B;
    const uint64_t band_row_bytes = bp->row_bytes * npos;
    uint64_t avail = band_bytes;
    if (!avail) {
        if (ctx->budget) {
            avail = ctx->budget > ctx->used ? ctx->budget - ctx->used : 0;
        } else {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(CP_ERR_HIP, "hipMemGetInfo failed");
            avail = free_b / 4;
        }
    }
    avail = avail > kStackSlack ? avail - kStackSlack : 0;
    const uint64_t fit = avail / band_row_bytes;  // band stack rows that fit
    if (fit < (uint64_t)bp->halo + 1)
        return fail(CP_ERR_OUT_OF_HBM_BUDGET, "a band of one output row needs %llu bytes, %llu available",
                    (unsigned long long)(((uint64_t)bp->halo + 1) * band_row_bytes + kStackSlack),
                    (unsigned long long)(avail + kStackSlack));
    bp->bh = (uint32_t)std::min<uint64_t>(fit - (uint64_t)bp->halo, out_h);
    const uint32_t band_h = bp->bh + (uint32_t)bp->halo;
    cp_stack& band = bp->band;
    band.ctx = ctx;
    band.width = W;
    band.height = band_h;
    band.channels = C;
    band.bps = B;
    band.capacity = band.count = npos;
    band.alloc_bytes = (uint64_t)band_h * band_row_bytes + kStackSlack;
    int rc = ctx->charge(band.alloc_bytes);
    if (rc) return rc;
    if (hipMalloc((void**)&band.dev, band.alloc_bytes) != hipSuccess) {
        ctx->uncharge(band.alloc_bytes);
        return fail(CP_ERR_HIP, "hipMalloc of a %llu-byte band stack failed", (unsigned long long)band.alloc_bytes);
    }
    bp->allocated = true;
    // position t's rows start at band row cy_t - min cy of its slot
    bp->pos.resize(npos);
    for (uint32_t t = 0; t < npos; ++t)
        bp->pos[t] = (uint64_t)t * band_h * W * C + ((uint64_t)(cy[t] - bp->min_cy) * W + (uint64_t)cx[t]) * C;
    bp->max_pos = *std::max_element(bp->pos.begin(), bp->pos.end());
    return CP_OK;
}

void free_bands(cp_ctx* ctx, BandPlan* bp, hipStream_t s) {
    if (!bp->allocated) return;
    (void)hipStreamSynchronize(s);
    (void)hipFree(bp->band.dev);
    ctx->uncharge(bp->band.alloc_bytes);
    bp->allocated = false;
}

// Copies the rows band r0 needs of every position into the band stack.  The previous band's
// kernels read the band stack until they finish: copies and kernels share the stream, so
// the upload starts after them.
int upload_band(const BandPlan& bp, const void* const* host_frames, uint32_t r0, uint32_t rows, hipStream_t s) {
    const uint64_t first = (uint64_t)((int64_t)r0 + bp.min_cy);  // first input row of the band
    const uint64_t in_rows = rows + (uint64_t)bp.halo;
    for (uint32_t t = 0; t < bp.band.count; ++t) {
        const uint8_t* src = (const uint8_t*)host_frames[bp.slot[t]] + first * bp.row_bytes;
        if (hipMemcpyAsync(bp.band.dev + (uint64_t)t * bp.band.height * bp.row_bytes, src, in_rows * bp.row_bytes,
                           hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(CP_ERR_HIP, "band upload failed");
    }
    return CP_OK;
}

}  // namespace

extern "C" {

CP_API int cp_outlier_host(cp_ctx* ctx, const void* const* host_frames, uint32_t n_host_frames, uint32_t width,
                           uint32_t height, uint32_t channels, uint32_t bytes_per_sample, const cp_frames* sel,
                           const cp_outlier_params* p, const cp_outputs* out, uint64_t band_bytes) {
    int rc = check_host_args(ctx, host_frames, sel, n_host_frames, width, height, channels, bytes_per_sample);
    if (rc) return rc;
    if (!p || !out || !out->px) return fail(CP_ERR_INVALID_ARG, "null argument");
    cp_stack view = host_view(ctx, n_host_frames, width, height, channels, bytes_per_sample);
    OutlierPrep prep;
    if ((rc = prepare_outlier(&view, sel, p, &prep))) return rc;
    const uint32_t npos = prep.k.n, ow = prep.k.out_w, oh = prep.k.out_h;
    const uint32_t C = channels, B = bytes_per_sample;

    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    // outputs first: the band stack gets what is left
    const uint64_t npx = (uint64_t)ow * oh;
    if ((rc = ctx->ensure(ctx->out_px, npx * C * B))) return rc;
    if (out->mask && (rc = ctx->ensure(ctx->out_mask, npx * C))) return rc;
    if (out->stats && (rc = ctx->ensure(ctx->out_stats, npx * sizeof(cp_pixel_stats)))) return rc;
    if ((rc = ctx->ensure(ctx->warn, 256))) return rc;
    BandPlan bp;
    if ((rc = plan_bands(ctx, host_frames, sel, width, C, B, npos, oh, band_bytes, &bp))) return rc;

    OutlierK k = prep.k;
    k.stack = bp.band.dev;
    k.stack_w = width;
    if ((rc = stage_outlier(ctx->arena, prep, bp.pos, s, &k))) {
        free_bands(ctx, &bp, s);
        return rc;
    }
    k.warnings = (unsigned long long*)ctx->warn.ptr;
    if (hipMemsetAsync(ctx->warn.ptr, 0, 8, s) != hipSuccess) {
        free_bands(ctx, &bp, s);
        return fail(CP_ERR_HIP, "hipMemsetAsync failed");
    }
    for (uint32_t r0 = 0; r0 < oh; r0 += bp.bh) {
        const uint32_t rows = std::min(bp.bh, oh - r0);
        if ((rc = upload_band(bp, host_frames, r0, rows, s))) {
            free_bands(ctx, &bp, s);
            return rc;
        }
        OutlierK kb = k;
        kb.out_h = rows;
        kb.pix_base = (uint64_t)r0 * ow;
        kb.out_px = (uint8_t*)ctx->out_px.ptr + (uint64_t)r0 * ow * C * B;
        kb.out_mask = out->mask ? (uint8_t*)ctx->out_mask.ptr + (uint64_t)r0 * ow * C : nullptr;
        kb.stats = out->stats ? (PixelStats*)ctx->out_stats.ptr + (uint64_t)r0 * ow : nullptr;
        if ((rc = launch_outlier(ctx, &bp.band, kb, bp.max_pos, 0, s))) {
            free_bands(ctx, &bp, s);
            return rc;
        }
    }
    unsigned long long w = 0;
    bool ok = hipMemcpyAsync(out->px, ctx->out_px.ptr, npx * C * B, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (out->mask)
        ok = ok && hipMemcpyAsync(out->mask, ctx->out_mask.ptr, npx * C, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (out->stats)
        ok = ok && hipMemcpyAsync(out->stats, ctx->out_stats.ptr, npx * sizeof(cp_pixel_stats), hipMemcpyDeviceToHost,
                                  s) == hipSuccess;
    ok = ok && hipMemcpyAsync(&w, ctx->warn.ptr, 8, hipMemcpyDeviceToHost, s) == hipSuccess;
    free_bands(ctx, &bp, s);
    if (!ok) return fail(CP_ERR_HIP, "output download failed");
    if (out->warnings) *out->warnings = w;
    return CP_OK;
}

CP_API int cp_simple_host(cp_ctx* ctx, const void* const* host_frames, uint32_t n_host_frames, uint32_t width,
                          uint32_t height, uint32_t channels, uint32_t bytes_per_sample, const cp_frames* sel,
                          const float weights[4], const cp_fade* fade, int32_t darker, void* out_px,
                          uint64_t band_bytes) {
    int rc = check_host_args(ctx, host_frames, sel, n_host_frames, width, height, channels, bytes_per_sample);
    if (rc) return rc;
    if (!weights || !out_px) return fail(CP_ERR_INVALID_ARG, "null argument");
    if ((rc = check_weights(weights))) return rc;
    cp_stack view = host_view(ctx, n_host_frames, width, height, channels, bytes_per_sample);
    Geometry g;
    if ((rc = resolve_frames(&view, sel, &g))) return rc;
    const uint32_t C = channels, B = bytes_per_sample;

    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const uint64_t px_bytes = (uint64_t)g.out_w * g.out_h * C * B;
    if ((rc = ctx->ensure(ctx->out_px, px_bytes))) return rc;
    BandPlan bp;
    if ((rc = plan_bands(ctx, host_frames, sel, width, C, B, g.n, g.out_h, band_bytes, &bp))) return rc;
    SimpleK k{};
    k.stack = bp.band.dev;
    k.pos_off = (const uint64_t*)ctx->arena.put(bp.pos.data(), 8 * g.n, s);
    k.n = g.n;
    k.stack_w = width;
    k.out_w = g.out_w;
    k.frame_offset = sel->frame_offset;
    k.darker = darker ? 1 : 0;
    for (int i = 0; i < 4; ++i) k.w[i] = weights[i];
    k.fade = fade_to_device(ctx, fade, &rc);
    if (rc || !k.pos_off) {
        free_bands(ctx, &bp, s);
        return rc ? rc : fail(CP_ERR_HIP, "parameter staging failed");
    }
    for (uint32_t r0 = 0; r0 < g.out_h; r0 += bp.bh) {
        const uint32_t rows = std::min(bp.bh, g.out_h - r0);
        if ((rc = upload_band(bp, host_frames, r0, rows, s))) {
            free_bands(ctx, &bp, s);
            return rc;
        }
        SimpleK kb = k;
        kb.out_h = rows;
        kb.out_px = (uint8_t*)ctx->out_px.ptr + (uint64_t)r0 * g.out_w * C * B;
        if ((rc = launch_simple(ctx, &bp.band, kb, s))) {
            free_bands(ctx, &bp, s);
            return rc;
        }
    }
    const bool ok = hipMemcpyAsync(out_px, ctx->out_px.ptr, px_bytes, hipMemcpyDeviceToHost, s) == hipSuccess;
    free_bands(ctx, &bp, s);
    if (!ok) return fail(CP_ERR_HIP, "output download failed");
    return CP_OK;
}

}  // extern "C"
