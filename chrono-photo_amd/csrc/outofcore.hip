// outofcore.hip -- cp_outlier_host: outlier composites of frame stacks that stay in host
// memory, streamed through the device in row bands.
//
// The reference never holds a whole stack in memory: its slicer writes time slices to disk
// and OutlierProcessor::process composites them one after another (src/chrono.rs:96-205,
// src/slicer.rs, src/streams.rs:115-219).  Here a band of output rows [r0, r0 + rows) needs,
// from every frame position t, input rows [r0 + cy_t, r0 + rows + cy_t) (cy_t: the position's
// crop origin); the band stack holds rows [r0 + min cy, r0 + rows + max cy) of each position,
// so one frame-offset table serves every band and the band is composited by the same kernels
// as a resident stack.  The random background draws by image pixel index (OutlierK::pix_base),
// so a band produces exactly the pixels the whole-image composite would.
#include <algorithm>
#include <vector>

#include "runtime.h"

using namespace cp;

extern "C" {

CP_API int cp_outlier_host(cp_ctx* ctx, const void* const* host_frames, uint32_t n_host_frames, uint32_t width,
                           uint32_t height, uint32_t channels, uint32_t bytes_per_sample, const cp_frames* sel,
                           const cp_outlier_params* p, const cp_outputs* out, uint64_t band_bytes) {
    if (!ctx || !host_frames || !sel || !p || !out || !out->px) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (width == 0 || height == 0 || n_host_frames == 0) return fail(CP_ERR_INVALID_ARG, "empty stack geometry");
    if (channels < 1 || channels > 4) return fail(CP_ERR_INVALID_ARG, "channels must be 1..4");
    if (bytes_per_sample != 1 && bytes_per_sample != 2)
        return fail(CP_ERR_INVALID_ARG, "bytes_per_sample must be 1 or 2");
    // the host frames seen as a stack: geometry, frame and crop checks as for cp_outlier
    cp_stack view;
    view.ctx = ctx;
    view.width = width;
    view.height = height;
    view.channels = channels;
    view.bps = bytes_per_sample;
    view.capacity = view.count = n_host_frames;
    OutlierPrep prep;
    int rc = prepare_outlier(&view, sel, p, &prep);
    if (rc) return rc;
    const uint32_t npos = prep.k.n, ow = prep.k.out_w, oh = prep.k.out_h;
    const uint32_t W = width, C = channels, B = bytes_per_sample;
    std::vector<uint32_t> slot(npos);
    std::vector<int64_t> cx(npos, 0), cy(npos, 0);
    for (uint32_t t = 0; t < npos; ++t) {
        slot[t] = sel->frames ? sel->frames[t] : t;
        if (!host_frames[slot[t]]) return fail(CP_ERR_INVALID_ARG, "host frame %u is null", slot[t]);
        if (sel->crop_xy) {
            cx[t] = sel->crop_xy[2 * t];
            cy[t] = sel->crop_xy[2 * t + 1];
        }
    }
    const int64_t min_cy = *std::min_element(cy.begin(), cy.end());
    const int64_t halo = *std::max_element(cy.begin(), cy.end()) - min_cy;
    const uint64_t row_bytes = (uint64_t)W * C * B;  // one row of one frame
    const uint64_t band_row_bytes = row_bytes * npos;

    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    // outputs first: the band stack gets what is left
    const uint64_t npx = (uint64_t)ow * oh;
    if ((rc = ctx->ensure(ctx->out_px, npx * C * B))) return rc;
    if (out->mask && (rc = ctx->ensure(ctx->out_mask, npx * C))) return rc;
    if (out->stats && (rc = ctx->ensure(ctx->out_stats, npx * sizeof(cp_pixel_stats)))) return rc;
    if ((rc = ctx->ensure(ctx->warn, 256))) return rc;
    uint64_t avail = band_bytes;
    if (!avail) {
        if (ctx->budget) {
            avail = ctx->budget > ctx->used ? ctx->budget - ctx->used : 0;
        } else {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
                return fail(CP_ERR_HIP, "hipMemGetInfo failed");
            avail = free_b / 4;
        }
    }
    avail = avail > kStackSlack ? avail - kStackSlack : 0;
    const uint64_t fit = avail / band_row_bytes;  // band stack rows that fit
    if (fit < (uint64_t)halo + 1)
        return fail(CP_ERR_OUT_OF_HBM_BUDGET, "a band of one output row needs %llu bytes, %llu available",
                    (unsigned long long)(((uint64_t)halo + 1) * band_row_bytes + kStackSlack),
                    (unsigned long long)(avail + kStackSlack));
    const uint32_t bh = (uint32_t)std::min<uint64_t>(fit - (uint64_t)halo, oh);  // output rows per band
    const uint32_t band_h = bh + (uint32_t)halo;                                 // stack rows per band

    cp_stack band;
    band.ctx = ctx;
    band.width = W;
    band.height = band_h;
    band.channels = C;
    band.bps = B;
    band.capacity = band.count = npos;
    band.alloc_bytes = (uint64_t)band_h * band_row_bytes + kStackSlack;
    if ((rc = ctx->charge(band.alloc_bytes))) return rc;
    if (hipMalloc((void**)&band.dev, band.alloc_bytes) != hipSuccess) {
        ctx->uncharge(band.alloc_bytes);
        return fail(CP_ERR_HIP, "hipMalloc of a %llu-byte band stack failed", (unsigned long long)band.alloc_bytes);
    }
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        (void)hipFree(band.dev);
        ctx->uncharge(band.alloc_bytes);
    };

    // one frame-offset table for every band: position t's rows start at band row cy_t - min cy
    std::vector<uint64_t> pos(npos);
    for (uint32_t t = 0; t < npos; ++t)
        pos[t] = (uint64_t)t * band_h * W * C + ((uint64_t)(cy[t] - min_cy) * W + (uint64_t)cx[t]) * C;
    const uint64_t max_pos = *std::max_element(pos.begin(), pos.end());
    OutlierK k = prep.k;
    k.stack = band.dev;
    k.stack_w = W;
    k.pos_off = (const uint64_t*)ctx->arena.put(pos.data(), 8 * npos, s);
    if (!prep.sample_mask.empty())
        k.sample_mask = (const uint32_t*)ctx->arena.put(prep.sample_mask.data(), 4 * prep.sample_mask.size(), s);
    if (!prep.fade_values.empty())
        k.fade.values = (const float*)ctx->arena.put(prep.fade_values.data(), 4 * prep.fade_values.size(), s);
    if (!k.pos_off || (!prep.sample_mask.empty() && !k.sample_mask) || (!prep.fade_values.empty() && !k.fade.values)) {
        release();
        return fail(CP_ERR_HIP, "parameter staging failed");
    }
    k.warnings = (unsigned long long*)ctx->warn.ptr;
    if (hipMemsetAsync(ctx->warn.ptr, 0, 8, s) != hipSuccess) {
        release();
        return fail(CP_ERR_HIP, "hipMemsetAsync failed");
    }

    for (uint32_t r0 = 0; r0 < oh; r0 += bh) {
        const uint32_t rows = std::min(bh, oh - r0);
        const uint64_t first = (uint64_t)(r0 + min_cy);  // first input row of the band
        const uint64_t in_rows = rows + (uint64_t)halo;
        // the kernels of the previous band read the band stack until they finish: copies and
        // kernels share the stream, so each upload starts after them
        for (uint32_t t = 0; t < npos; ++t) {
            const uint8_t* src = (const uint8_t*)host_frames[slot[t]] + first * row_bytes;
            if (hipMemcpyAsync(band.dev + (uint64_t)t * band_h * row_bytes, src, in_rows * row_bytes,
                               hipMemcpyHostToDevice, s) != hipSuccess) {
                release();
                return fail(CP_ERR_HIP, "band upload failed");
            }
        }
        OutlierK kb = k;
        kb.out_h = rows;
        kb.pix_base = (uint64_t)r0 * ow;
        kb.out_px = (uint8_t*)ctx->out_px.ptr + (uint64_t)r0 * ow * C * B;
        kb.out_mask = out->mask ? (uint8_t*)ctx->out_mask.ptr + (uint64_t)r0 * ow * C : nullptr;
        kb.stats = out->stats ? (PixelStats*)ctx->out_stats.ptr + (uint64_t)r0 * ow : nullptr;
        if ((rc = launch_outlier(ctx, &band, kb, max_pos, 0, s))) {
            release();
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
    release();
    if (!ok) return fail(CP_ERR_HIP, "output download failed");
    if (out->warnings) *out->warnings = w;
    return CP_OK;
}

}  // extern "C"
