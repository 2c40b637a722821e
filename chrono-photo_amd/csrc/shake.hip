// shake.hip -- cp_shake: ShakeAnalyzer::analyze (src/shake.rs:190-313) on a resident stack,
// and cp_shake_noalloc, the same search with its scratch in a static device array.
#include "k_shake.h"
#include "runtime.h"

using namespace cp;

namespace {

// Scratch of cp_shake_noalloc: the per-frame candidate sums ([n][(2s+1)^2] + 1 u64) and, for
// host outputs, the offsets staged before their copy.  One array per device, for the life of
// the process; its device address is looked up once per device when a context is created on
// it (shake_scratch_init) and only that address is used on the host.
constexpr uint64_t kShakeScratchBytes = 8ull << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned long long scratch[kShakeScratchBytes / 8];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Clears the sum table with a kernel (a memset node would misbehave if a caller captured the
// call in a graph).
__global__ void __launch_bounds__(256) zero_u64_kernel(unsigned long long* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

// Offsets of a search without anchors or of a single frame: (0, 0) for position 0 and
// (-s, -s) for the others (every sum is 0 and min_by_key keeps index 0).
__global__ void __launch_bounds__(256) shake_trivial_kernel(int32_t* offsets, uint32_t n, int32_t s) {
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) {
        offsets[2 * t] = t == 0 ? 0 : -s;
        offsets[2 * t + 1] = t == 0 ? 0 : -s;
    }
}

}  // namespace

int cp::shake_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CP_ERR_INVALID_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return CP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CP_ERR_HIP, "scratch address lookup on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return CP_OK;
}

// The search of cp_shake.  scratch_dev == nullptr: the sum table comes from hipMalloc; else it
// lives in scratch_dev (kShakeScratchBytes).  dev_out: offsets is a device array written
// asynchronously on the context stream; else a host array and the call waits.
static int shake_run(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                     const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                     uint32_t search_radius, int32_t* offsets, bool dev_out, void* scratch_dev) {
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
    const int32_t ssize = (int32_t)(2 * s + 1);
    const uint64_t noff = (uint64_t)ssize * ssize;
    const uint64_t sums_bytes = (noff * n + 1) * 8;  // per-frame tables + the window constant
    const uint64_t sums_al = (sums_bytes + 255) & ~uint64_t(255);
    if (scratch_dev && sums_al + (dev_out ? 0 : 8ull * n) > kShakeScratchBytes)
        return fail(CP_ERR_CAPACITY,
                    "cp_shake_noalloc: %u frames x %llu candidate offsets need %llu bytes of scratch (at most %llu)",
                    n, (unsigned long long)noff, (unsigned long long)(sums_al + (dev_out ? 0 : 8ull * n)),
                    (unsigned long long)kShakeScratchBytes);
    if (n == 1 || n_anchors == 0) {
        // No anchors: every diff is 0 and min_by_key keeps index 0, i.e. offset (-s, -s).
        if (dev_out) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            (void)hipSetDevice(ctx->device);
            hipLaunchKernelGGL(shake_trivial_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, offsets, n,
                               (int32_t)s);
            CP_HIP_TRY(hipGetLastError());
            return CP_OK;
        }
        offsets[0] = 0;
        offsets[1] = 0;
        for (uint32_t t = 1; t < n; ++t) { offsets[2 * t] = -(int32_t)s; offsets[2 * t + 1] = -(int32_t)s; }
        return CP_OK;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t hs = ctx->stream;
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
    int rc = CP_OK;
    if (scratch_dev) {
        sums.ptr = scratch_dev;
        offs.ptr = dev_out ? (void*)offsets : (void*)((uint8_t*)scratch_dev + sums_al);
    } else {
        const uint64_t offs_bytes = dev_out ? 0 : 8ull * n;
        if ((rc = ctx->charge(sums_bytes + offs_bytes))) return rc;
        if (hipMalloc(&sums.ptr, sums_bytes) != hipSuccess ||
            (!dev_out && hipMalloc(&offs.ptr, offs_bytes) != hipSuccess)) {
            if (sums.ptr) (void)hipFree(sums.ptr);
            ctx->uncharge(sums_bytes + offs_bytes);
            return fail(CP_ERR_HIP, "hipMalloc for shake sums failed");
        }
        if (dev_out) offs.ptr = offsets;
    }
    k.sums = (unsigned long long*)sums.ptr;
    k.offsets = (int32_t*)offs.ptr;
    // every sum starts at zero: the previous call's contents (a shared scratch) are cleared
    const uint64_t n_sums = sums_bytes / 8;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((uint32_t)std::min<uint64_t>(1024, (n_sums + 255) / 256)), dim3(256), 0,
                       hs, k.sums, n_sums);
    hipError_t e = hipGetLastError();
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
        if (e == hipSuccess && !dev_out) e = hipMemcpyAsync(offsets, offs.ptr, n * 8, hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess && !dev_out) e = hipStreamSynchronize(hs);
        if (e != hipSuccess && rc == CP_OK) rc = fail(CP_ERR_HIP, "shake kernels: %s", hipGetErrorString(e));
    }
    if (!scratch_dev) {
        (void)hipStreamSynchronize(hs);
        (void)hipFree(sums.ptr);
        if (!dev_out) (void)hipFree(offs.ptr);
        ctx->uncharge(sums_bytes + (dev_out ? 0 : 8ull * n));
    }
    return rc;
}

extern "C" CP_API int cp_shake(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                               const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                               uint32_t search_radius, int32_t* offsets) {
    return shake_run(ctx, st, frames, n, anchors_xy, n_anchors, anchor_radius, search_radius, offsets, false,
                     nullptr);
}

extern "C" CP_API int cp_shake_noalloc(cp_ctx* ctx, const cp_stack* st, const uint32_t* frames, uint32_t n,
                                       const int32_t* anchors_xy, uint32_t n_anchors, uint32_t anchor_radius,
                                       uint32_t search_radius, int32_t* offsets, uint32_t flags) {
    if (!ctx) return fail(CP_ERR_INVALID_ARG, "null ctx");
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) p = g_scratch_ptr[ctx->device];
    }
    if (!p) return fail(CP_ERR_HIP, "no scratch address for device %d", ctx->device);
    return shake_run(ctx, st, frames, n, anchors_xy, n_anchors, anchor_radius, search_radius, offsets,
                     (flags & CP_FLAG_DEVICE_OUT) != 0, p);
}
