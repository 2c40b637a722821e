// shake.hip -- cp_shake: ShakeAnalyzer::analyze (src/shake.rs:190-313) on a resident stack.
#include "k_shake.h"
#include "runtime.h"

using namespace cp;

extern "C" CP_API int cp_shake(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                               const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                               uint32_t search_radius, int32_t* offsets) {
    if (!ctx || !st || !offsets || (!anchors_xy && n_anchors)) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (st->ctx != ctx) return fail(CP_ERR_INVALID_ARG, "stack belongs to another context");
    if (n == 0) return fail(CP_ERR_INVALID_ARG, "shake analysis needs at least one frame");
    if (anchor_radius > 4096 || search_radius > 4096) return fail(CP_ERR_INVALID_ARG, "radius too large");
    const int64_t r = anchor_radius, s = search_radius, W = st->width, H = st->height;
    // fill_windows (:328-341) then calc_diffs (:363-381) index the image; out-of-range
    // coordinates panic in the reference.
    for (uint32_t a = 0; a < n_anchors; ++a) {
        int64_t cx = anchors_xy[2 * a], cy = anchors_xy[2 * a + 1];
        if (cx - r < 0 || cy - r < 0 || cx + r >= W || cy + r >= H)
            return fail(CP_ERR_ANCHOR_OUT_OF_RANGE, "Image coordinate out of range: anchor (%lld, %lld) radius %lld",
                        (long long)cx, (long long)cy, (long long)r);
        if (n > 1 && (cx - r - s < 0 || cy - r - s < 0 || cx + r + s >= W || cy + r + s >= H))
            return fail(CP_ERR_ANCHOR_OUT_OF_RANGE,
                        "Image coordinate out of range: anchor (%lld, %lld) search radius %lld",
                        (long long)cx, (long long)cy, (long long)s);
    }
    std::vector<uint64_t> pos(n);
    for (uint32_t t = 0; t < n; ++t) {
        uint32_t slot = frames ? frames[t] : t;
        if (slot >= st->count) return fail(CP_ERR_INVALID_ARG, "frame slot %u not loaded", slot);
        pos[t] = (uint64_t)slot * st->frame_elems();
    }
    offsets[0] = 0;
    offsets[1] = 0;
    if (n == 1 || n_anchors == 0) {
        // No anchors: every diff is 0 and min_by_key keeps index 0, i.e. offset (-s, -s).
        for (uint32_t t = 1; t < n; ++t) { offsets[2 * t] = -(int32_t)s; offsets[2 * t + 1] = -(int32_t)s; }
        return CP_OK;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t hs = ctx->stream;
    const int32_t ssize = (int32_t)(2 * s + 1);
    const uint64_t noff = (uint64_t)ssize * ssize;
    ShakeK k{};
    k.stack = st->dev;
    k.pos_off = (const uint64_t*)ctx->arena.put(pos.data(), 8 * n, hs);
    k.anchors = (const int32_t*)ctx->arena.put(anchors_xy, 8 * n_anchors, hs);
    if (!k.pos_off || !k.anchors) return fail(CP_ERR_HIP, "parameter staging failed");
    k.stack_w = st->width;
    k.n = n;
    k.n_anchors = n_anchors;
    k.r = (int32_t)r;
    k.s = (int32_t)s;
    DeviceBuffer sums, offs;
    const uint64_t sums_bytes = (noff * n + 1) * 8;  // per-frame tables + the window constant
    int rc = ctx->charge(sums_bytes + n * 8);
    if (rc) return rc;
    if (hipMalloc(&sums.ptr, sums_bytes) != hipSuccess || hipMalloc(&offs.ptr, n * 8) != hipSuccess) {
        if (sums.ptr) (void)hipFree(sums.ptr);
        ctx->uncharge(sums_bytes + n * 8);
        return fail(CP_ERR_HIP, "hipMalloc for shake sums failed");
    }
    k.sums = (unsigned long long*)sums.ptr;
    k.offsets = (int32_t*)offs.ptr;
    hipError_t e = hipMemsetAsync(sums.ptr, 0, sums_bytes, hs);
    // Launch plan: strips of kShakeStrip offsets per lane, whole waves per row sub-chunk,
    // about 8 window rows per sub-chunk and all candidate rows per workgroup; where that
    // does not fit LDS, the larger of the two row counts shrinks.
    constexpr int32_t K = kShakeStrip;
    const int32_t size = (int32_t)(2 * r + 1), C = (int32_t)st->channels;
    k.spr = (ssize + K - 1) / K;
    k.pitch_px = k.spr * K + (int32_t)(2 * r);
    int32_t block = 256;
    auto set_lanes = [&](int32_t orows) {
        k.strips_pad = (k.spr * orows + 63) / 64 * 64;
        k.nsub = k.strips_pad <= 256 ? 256 / k.strips_pad : 1;
        block = k.strips_pad <= 256 ? k.strips_pad * k.nsub : 256;
    };
    // Exact f64 sums: each lane adds ceil(rows / nsub) * size * C products below maxsq.
    const double maxsq = st->bps == 2 ? 65535.0 * 65535.0 : 255.0 * 255.0;
    auto fits = [&](int32_t rows, int32_t orows) {
        set_lanes(orows);
        const double terms = (double)((rows + k.nsub - 1) / k.nsub) * size * C;
        return shake_lds(rows, orows, (int32_t)r, (int32_t)s, k.pitch_px, C).bytes <= 64 * 1024 &&
               terms * maxsq < 9007199254740992.0;
    };
    int32_t orows = ssize;
    set_lanes(orows);
    const int32_t want = std::min<int32_t>(size, 8 * k.nsub);
    const int32_t even_chunks = (size + want - 1) / want;
    int32_t rows = (size + even_chunks - 1) / even_chunks;
    while (!fits(rows, orows) && (rows > 1 || orows > 1)) {
        if (orows > rows) --orows;
        else --rows;
    }
    const bool ok = fits(rows, orows);  // also leaves the lane split of the final orows in k
    k.chunk_rows = rows;
    k.cand_rows = orows;
    k.cand_chunks = (ssize + orows - 1) / orows;
    const uint32_t nchunk = (uint32_t)((size + rows - 1) / rows);
    const size_t lds = shake_lds(rows, orows, (int32_t)r, (int32_t)s, k.pitch_px, C).bytes;
    if (e == hipSuccess && !ok) {
        e = hipErrorInvalidValue;
        rc = fail(CP_ERR_INVALID_ARG,
                  "anchor search needs %zu bytes of LDS for one window row and one candidate row (64 KiB at most)",
                  (size_t)shake_lds(1, 1, (int32_t)r, (int32_t)s, k.pitch_px, C).bytes);
    } else if (e == hipSuccess && (uint64_t)n_anchors * k.cand_chunks > 65535) {
        e = hipErrorInvalidValue;
        rc = fail(CP_ERR_INVALID_ARG, "%u anchors x %d candidate-row chunks exceed 65535 workgroup rows",
                  n_anchors, k.cand_chunks);
    }
    if (e == hipSuccess) {
        dim3 grid(n - 1, n_anchors * (uint32_t)k.cand_chunks, nchunk), blk(block);
        auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, grid, blk, lds, hs, k); };
        const bool wide = st->bps == 2;
        switch (st->channels) {
        case 1: wide ? launch(shake_ssd_kernel<2, 1>) : launch(shake_ssd_kernel<1, 1>); break;
        case 2: wide ? launch(shake_ssd_kernel<2, 2>) : launch(shake_ssd_kernel<1, 2>); break;
        case 3: wide ? launch(shake_ssd_kernel<2, 3>) : launch(shake_ssd_kernel<1, 3>); break;
        default: wide ? launch(shake_ssd_kernel<2, 4>) : launch(shake_ssd_kernel<1, 4>); break;
        }
        if (wide) hipLaunchKernelGGL(shake_argmin_kernel<2>, dim3(n), dim3(256), 0, hs, k);
        else hipLaunchKernelGGL(shake_argmin_kernel<1>, dim3(n), dim3(256), 0, hs, k);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(offsets, offs.ptr, n * 8, hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
        if (e != hipSuccess && rc == CP_OK) rc = fail(CP_ERR_HIP, "shake kernels: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(hs);
    (void)hipFree(sums.ptr);
    (void)hipFree(offs.ptr);
    ctx->uncharge(sums_bytes + n * 8);
    return rc;
}
