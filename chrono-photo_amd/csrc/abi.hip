// abi.hip -- C ABI implementation: runtime, stacks, option values, compositing entry
// points.  The contract of each entry point is documented in include/chrono_hip.h.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include "k_outlier_generic.h"
#include "kernels_fast.h"
#include "runtime.h"

namespace cp {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int ParamArena::init(size_t bytes) {
    if (hipHostMalloc((void**)&host, bytes, hipHostMallocDefault) != hipSuccess) return -1;
    if (hipMalloc((void**)&dev, bytes) != hipSuccess) {
        (void)hipHostFree(host);
        host = nullptr;
        return -1;
    }
    cap = bytes;
    head = 0;
    return 0;
}

void ParamArena::release() {
    if (host) (void)hipHostFree(host);
    if (dev) (void)hipFree(dev);
    host = dev = nullptr;
    cap = head = 0;
}

void* ParamArena::put(const void* data, size_t bytes, hipStream_t s) {
    if (bytes == 0) bytes = 1;
    size_t need = (bytes + 255) & ~size_t(255);
    if (need > cap) return nullptr;
    if (head + need > cap) {
        if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
        head = 0;
    }
    uint8_t* h = host + head;
    uint8_t* d = dev + head;
    if (data) memcpy(h, data, bytes);
    else memset(h, 0, bytes);
    if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return nullptr;
    head += need;
    return d;
}

int resolve_frames(const cp_stack* st, const cp_frames* fr, Geometry* g) {
    if (!st || !fr) return fail(CP_ERR_INVALID_ARG, "null stack or frames");
    if (fr->n == 0) return fail(CP_ERR_INVALID_ARG, "a composite needs at least one frame");
    g->n = fr->n;
    g->out_w = fr->out_width ? fr->out_width : st->width;
    g->out_h = fr->out_height ? fr->out_height : st->height;
    if (g->out_w > st->width || g->out_h > st->height)
        return fail(CP_ERR_LAYOUT_MISMATCH, "output %ux%u larger than stack frames %ux%u", g->out_w,
                    g->out_h, st->width, st->height);
    g->pos_off.resize(fr->n);
    for (uint32_t t = 0; t < fr->n; ++t) {
        uint32_t slot = fr->frames ? fr->frames[t] : t;
        if (slot >= st->count)
            return fail(CP_ERR_INVALID_ARG, "frame slot %u not loaded (stack holds %u)", slot,
                        st->count);
        int64_t cx = 0, cy = 0;
        if (fr->crop_xy) {
            cx = fr->crop_xy[2 * t];
            cy = fr->crop_xy[2 * t + 1];
        }
        if (cx < 0 || cy < 0 || cx + g->out_w > st->width || cy + g->out_h > st->height)
            return fail(CP_ERR_LAYOUT_MISMATCH, "crop of frame %u at (%lld,%lld) size %ux%u leaves the frame",
                        t, (long long)cx, (long long)cy, g->out_w, g->out_h);
        g->pos_off[t] = (uint64_t)slot * st->frame_elems() +
                        ((uint64_t)cy * st->width + (uint64_t)cx) * st->channels;
    }
    return CP_OK;
}

FadeK fade_to_device(cp_ctx* ctx, const cp_fade* f, int* rc) {
    FadeK k{};
    *rc = CP_OK;
    if (!f || f->is_none) {
        k.is_none = 1;
        return k;
    }
    if (f->len == 0 || !f->values) {
        *rc = fail(CP_ERR_INVALID_ARG, "fade table is empty");
        return k;
    }
    k.is_none = 0;
    k.repeat = f->repeat;
    k.absolute = f->absolute;
    k.offset = f->offset;
    k.len = f->len;
    k.values = (const float*)ctx->arena.put(f->values, sizeof(float) * f->len, ctx->stream);
    if (!k.values) *rc = fail(CP_ERR_HIP, "parameter staging failed for the fade table");
    return k;
}

int check_weights(const float w[4]) {
    for (int i = 0; i < 4; ++i)
        if (std::isnan(w[i]) || std::isinf(w[i])) return fail(CP_ERR_INVALID_ARG, "weights must be finite");
    return CP_OK;
}

int prepare_outlier(const cp_stack* st, const cp_frames* frames, const cp_outlier_params* p, OutlierPrep* out) {
    if (p->background < 0 || p->background > 3) return fail(CP_ERR_INVALID_ARG, "bad background mode");
    if (p->outlier < 0 || p->outlier > 5) return fail(CP_ERR_INVALID_ARG, "bad outlier mode");
    int rc = check_weights(p->weights);
    if (rc) return rc;
    Geometry g;
    if ((rc = resolve_frames(st, frames, &g))) return rc;
    const uint32_t S = p->sample_pos ? p->n_sample : g.n;
    if (S == 0) return fail(CP_ERR_INVALID_ARG, "empty median sample");
    if (!p->threshold.absolute && S < 3)
        return fail(CP_ERR_TOO_FEW_FRAMES_REL, "relative thresholds need at least 3 samples (got %u)", S);
    if (p->sample_pos) {
        out->sample_mask.assign((g.n + 31) / 32, 0u);
        uint32_t prev = 0;
        for (uint32_t i = 0; i < p->n_sample; ++i) {
            const uint32_t t = p->sample_pos[i];
            if (t >= g.n || (i > 0 && t <= prev))
                return fail(CP_ERR_INVALID_ARG, "sample positions must be sorted, unique and < n");
            out->sample_mask[t >> 5] |= 1u << (t & 31);
            prev = t;
        }
        out->sample_pos.assign(p->sample_pos, p->sample_pos + p->n_sample);
    }
    if (p->fade.is_none == 0) {
        if (p->fade.len == 0 || !p->fade.values) return fail(CP_ERR_INVALID_ARG, "fade table is empty");
        out->fade_values.assign(p->fade.values, p->fade.values + p->fade.len);
    }
    OutlierK& k = out->k;
    k = OutlierK{};
    k.stack = st->dev;
    k.n = g.n;
    k.n_sample = S;
    k.stack_w = st->width;
    k.out_w = g.out_w;
    k.out_h = g.out_h;
    k.frame_offset = frames->frame_offset;
    k.th = ThresholdK{p->threshold.absolute, p->threshold.min, p->threshold.max, p->threshold.scale};
    k.background = p->background;
    k.outlier = p->outlier;
    for (int i = 0; i < 4; ++i) k.w[i] = p->weights[i];
    k.fade.is_none = p->fade.is_none ? 1 : 0;
    k.fade.repeat = p->fade.repeat;
    k.fade.absolute = p->fade.absolute;
    k.fade.offset = p->fade.offset;
    k.fade.len = p->fade.len;
    k.seed = p->seed;
    out->pos_off = std::move(g.pos_off);
    return CP_OK;
}

int stage_outlier(ParamArena& arena, const OutlierPrep& prep, const std::vector<uint64_t>& pos, hipStream_t s,
                  OutlierK* k) {
    k->pos_off = (const uint64_t*)arena.put(pos.data(), 8 * pos.size(), s);
    if (!k->pos_off) return fail(CP_ERR_HIP, "parameter staging failed");
    if (!prep.sample_mask.empty()) {
        k->sample_mask = (const uint32_t*)arena.put(prep.sample_mask.data(), 4 * prep.sample_mask.size(), s);
        std::vector<uint64_t> so(prep.sample_pos.size());
        for (size_t i = 0; i < so.size(); ++i) so[i] = pos[prep.sample_pos[i]];
        k->samp_off = (const uint64_t*)arena.put(so.data(), 8 * so.size(), s);
        if (!k->sample_mask || !k->samp_off) return fail(CP_ERR_HIP, "parameter staging failed");
    }
    if (!prep.fade_values.empty()) {
        k->fade.values = (const float*)arena.put(prep.fade_values.data(), 4 * prep.fade_values.size(), s);
        if (!k->fade.values) return fail(CP_ERR_HIP, "parameter staging failed");
    }
    return CP_OK;
}

}  // namespace cp

int cp_ctx::ensure(cp::DeviceBuffer& b, size_t bytes) {
    if (b.bytes >= bytes) return CP_OK;
    if (b.ptr) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(b.ptr);
        uncharge(b.bytes);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    int rc = charge(bytes);
    if (rc) return rc;
    if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
        uncharge(bytes);
        return cp::fail(CP_ERR_HIP, "hipMalloc(%zu) failed", bytes);
    }
    b.bytes = bytes;
    return CP_OK;
}

int cp_ctx::choice_for(uint32_t bps, uint32_t channels, uint32_t n, bool absolute) const {
    const int b = cp::bucket_for(n);
    if (b < 0 || bps < 1 || bps > 2 || channels < 1 || channels > 4) return -1;
    const int t = tuned[bps - 1][channels - 1][absolute ? 0 : 1][b];
    if (t >= 0) return t;
    const int rel = absolute ? -1 : cp::kRelDefault[bps - 1][channels >= 3][b];
    return rel >= 0 ? rel : cp::kBucketCand[bps - 1][channels >= 3][b][0];
}

int cp_ctx::charge(uint64_t bytes) {
    if (budget && used + bytes > budget)
        return cp::fail(CP_ERR_OUT_OF_HBM_BUDGET, "HBM budget %llu exceeded (%llu used, %llu asked)",
                        (unsigned long long)budget, (unsigned long long)used,
                        (unsigned long long)bytes);
    used += bytes;
    return CP_OK;
}

using namespace cp;

extern "C" {

CP_API const char* cp_last_error(void) { return g_err.c_str(); }
CP_API const char* cp_version(void) { return "chrono-photo-hip 0.1.0 (gfx950)"; }

CP_API int cp_ctx_create(int32_t device, uint64_t hbm_budget_bytes, cp_ctx** out) {
    if (!out) return fail(CP_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(CP_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(CP_ERR_INVALID_ARG, "device %d out of range", device);
    CP_HIP_TRY(hipSetDevice(device));
    if (int rc = shake_scratch_init(device)) return rc;
    cp_ctx* c = new cp_ctx();
    memset(c->tuned, -1, sizeof(c->tuned));
    c->device = device;
    c->budget = hbm_budget_bytes;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(CP_ERR_HIP, "stream creation failed");
    }
    if (c->arena.init(8u << 20) != 0) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(CP_ERR_HIP, "parameter arena allocation failed");
    }
    *out = c;
    return CP_OK;
}

CP_API int cp_ctx_destroy(cp_ctx* ctx) {
    if (!ctx) return CP_OK;
    (void)hipSetDevice(ctx->device);
    // Handles still alive: servers first (their workers launch on stacks), then batches, then
    // stacks.  Each keeps its host shell until its own destroy call, which then only frees it.
    std::vector<cp_server*> servers;
    std::vector<cp_batch*> batches;
    std::vector<cp_stack*> stacks;
    {
        std::lock_guard<std::mutex> l(ctx->reg_mu);
        servers.assign(ctx->servers.begin(), ctx->servers.end());
        batches.assign(ctx->batches.begin(), ctx->batches.end());
        stacks.assign(ctx->stacks.begin(), ctx->stacks.end());
        ctx->servers.clear();
        ctx->batches.clear();
        ctx->stacks.clear();
    }
    for (cp_server* v : servers) server_orphan(v);
    for (cp_batch* b : batches) batch_orphan(b);
    (void)hipStreamSynchronize(ctx->stream);
    for (cp_stack* st : stacks) {
        if (st->owned && st->dev) (void)hipFree(st->dev);
        st->dev = nullptr;
        st->count = st->capacity = 0;
        st->ctx = nullptr;
    }
    for (DeviceBuffer* b : {&ctx->out_px, &ctx->out_mask, &ctx->out_stats, &ctx->warn})
        if (b->ptr) (void)hipFree(b->ptr);
    ctx->arena.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CP_OK;
}

CP_API int cp_ctx_synchronize(cp_ctx* ctx) {
    if (!ctx) return fail(CP_ERR_INVALID_ARG, "null ctx");
    CP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CP_OK;
}

CP_API void* cp_ctx_stream(cp_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

CP_API int cp_stack_create(cp_ctx* ctx, uint32_t width, uint32_t height, uint32_t channels,
                           uint32_t bytes_per_sample, uint32_t capacity, cp_stack** out) {
    if (!ctx || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0 || capacity == 0)
        return fail(CP_ERR_INVALID_ARG, "empty stack geometry");
    if (channels < 1 || channels > 4) return fail(CP_ERR_INVALID_ARG, "channels must be 1..4");
    if (bytes_per_sample != 1 && bytes_per_sample != 2)
        return fail(CP_ERR_INVALID_ARG, "bytes_per_sample must be 1 or 2");
    cp_stack* s = new cp_stack();
    s->ctx = ctx;
    s->width = width;
    s->height = height;
    s->channels = channels;
    s->bps = bytes_per_sample;
    s->capacity = capacity;
    uint64_t bytes = s->frame_bytes() * capacity + kStackSlack;
    int rc = ctx->charge(bytes);
    if (rc) { delete s; return rc; }
    (void)hipSetDevice(ctx->device);
    if (hipMalloc((void**)&s->dev, bytes) != hipSuccess) {
        ctx->uncharge(bytes);
        delete s;
        return fail(CP_ERR_HIP, "hipMalloc of %llu stack bytes failed", (unsigned long long)bytes);
    }
    s->alloc_bytes = bytes;
    s->owned = true;
    ctx->attach(s);
    *out = s;
    return CP_OK;
}

CP_API int cp_stack_wrap(cp_ctx* ctx, void* device_ptr, uint32_t width, uint32_t height,
                         uint32_t channels, uint32_t bytes_per_sample, uint32_t n_frames,
                         cp_stack** out) {
    if (!ctx || !out || !device_ptr) return fail(CP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0 || n_frames == 0)
        return fail(CP_ERR_INVALID_ARG, "empty stack geometry");
    if (channels < 1 || channels > 4) return fail(CP_ERR_INVALID_ARG, "channels must be 1..4");
    if (bytes_per_sample != 1 && bytes_per_sample != 2)
        return fail(CP_ERR_INVALID_ARG, "bytes_per_sample must be 1 or 2");
    cp_stack* s = new cp_stack();
    s->ctx = ctx;
    s->width = width;
    s->height = height;
    s->channels = channels;
    s->bps = bytes_per_sample;
    s->capacity = s->count = n_frames;
    s->dev = (uint8_t*)device_ptr;
    s->alloc_bytes = s->frame_bytes() * n_frames;
    s->owned = false;
    ctx->attach(s);
    *out = s;
    return CP_OK;
}

static int check_live(const cp_stack* st) {
    if (!st) return fail(CP_ERR_INVALID_ARG, "null stack");
    if (!st->ctx) return fail(CP_ERR_INVALID_ARG, "the stack's context was destroyed");
    return CP_OK;
}

CP_API int cp_stack_upload(cp_stack* st, uint32_t slot, const void* host_frames, uint32_t n) {
    if (int rc = check_live(st)) return rc;
    if (!host_frames && n) return fail(CP_ERR_INVALID_ARG, "null argument");
    if ((uint64_t)slot + n > st->capacity)
        return fail(CP_ERR_CAPACITY, "slots [%u,%u) exceed capacity %u", slot, slot + n, st->capacity);
    (void)hipSetDevice(st->ctx->device);
    CP_HIP_TRY(hipMemcpyAsync(st->dev + (uint64_t)slot * st->frame_bytes(), host_frames,
                              (uint64_t)n * st->frame_bytes(), hipMemcpyHostToDevice,
                              st->ctx->stream));
    CP_HIP_TRY(hipStreamSynchronize(st->ctx->stream));
    if (slot + n > st->count) st->count = slot + n;
    return CP_OK;
}

CP_API int cp_stack_push(cp_stack* st, const void* host_frames, uint32_t n, uint32_t* first_slot) {
    if (int rc = check_live(st)) return rc;
    uint32_t first = st->count;
    int rc = cp_stack_upload(st, first, host_frames, n);
    if (rc) return rc;
    if (first_slot) *first_slot = first;
    return CP_OK;
}

CP_API int cp_stack_download(const cp_stack* st, uint32_t slot, void* host_frames, uint32_t n) {
    if (int rc = check_live(st)) return rc;
    if (!host_frames && n) return fail(CP_ERR_INVALID_ARG, "null argument");
    if ((uint64_t)slot + n > st->count)
        return fail(CP_ERR_INVALID_ARG, "slots [%u,%u) not loaded (stack holds %u)", slot, slot + n, st->count);
    (void)hipSetDevice(st->ctx->device);
    CP_HIP_TRY(hipMemcpyAsync(host_frames, st->dev + (uint64_t)slot * st->frame_bytes(),
                              (uint64_t)n * st->frame_bytes(), hipMemcpyDeviceToHost, st->ctx->stream));
    CP_HIP_TRY(hipStreamSynchronize(st->ctx->stream));
    return CP_OK;
}

CP_API int cp_stack_commit(cp_stack* st, uint32_t count) {
    if (int rc = check_live(st)) return rc;
    if (count > st->capacity) return fail(CP_ERR_CAPACITY, "count %u exceeds capacity %u", count, st->capacity);
    st->count = count;
    return CP_OK;
}

CP_API int cp_stack_info_get(const cp_stack* st, cp_stack_info* out) {
    if (!st || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    out->width = st->width;
    out->height = st->height;
    out->channels = st->channels;
    out->bytes_per_sample = st->bps;
    out->capacity = st->capacity;
    out->count = st->count;
    out->device_ptr = st->dev;
    return CP_OK;
}

CP_API int cp_stack_destroy(cp_stack* st) {
    if (!st) return CP_OK;
    if (!st->ctx) {  // its context was destroyed first and already released the memory
        delete st;
        return CP_OK;
    }
    st->ctx->detach(st);
    if (st->owned && st->dev) {
        (void)hipSetDevice(st->ctx->device);
        (void)hipStreamSynchronize(st->ctx->stream);
        (void)hipFree(st->dev);
        st->ctx->uncharge(st->alloc_bytes);
    }
    delete st;
    return CP_OK;
}

CP_API int cp_threshold_new(int32_t absolute, float min, float max, uint32_t bytes_per_sample,
                            cp_threshold* out) {
    if (!out) return fail(CP_ERR_INVALID_ARG, "null out");
    if (bytes_per_sample != 1 && bytes_per_sample != 2)
        return fail(CP_ERR_INVALID_ARG, "bytes_per_sample must be 1 or 2");
    const float range = bytes_per_sample == 1 ? 255.0f : 65535.0f;
    out->absolute = absolute ? 1 : 0;
    if (absolute) {
        out->min = min * range;
        out->max = max * range;
        out->scale = 1.0f / ((max - min) * range);
    } else {
        out->min = min;
        out->max = max;
        out->scale = 1.0f / (max - min);
    }
    return CP_OK;
}

CP_API float cp_threshold_blend_value(const cp_threshold* t, float dist) {
    if (dist <= t->min) return 0.0f;
    if (dist >= t->max) return 1.0f;
    return (dist - t->min) * t->scale;
}

CP_API int cp_fade_table(const int32_t* frames, const float* values, uint32_t nkeys, float* out,
                         uint32_t cap, uint32_t* len) {
    if (!frames || !values || !len) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (nkeys < 2) return fail(CP_ERR_INVALID_ARG, "Fade requires at least two frames specified.");
    for (uint32_t i = 1; i < nkeys; ++i)
        if (frames[i] < frames[i - 1])
            return fail(CP_ERR_INVALID_ARG, "fade keyframes must be ordered by frame");
    const int32_t offset = frames[0];
    const uint32_t n = (uint32_t)(frames[nkeys - 1] - offset);
    uint32_t idx = 0;
    for (uint32_t i = 0; i < n + 1; ++i) {
        int32_t f1 = frames[idx], f2 = frames[idx + 1];
        float v1 = values[idx], v2 = values[idx + 1];
        int32_t frame = (int32_t)i + offset;
        float v = v1 + (v2 - v1) * (float)(frame - f1) / (float)(f2 - f1);
        if (out && i < cap) out[i] = v;
        if (frame == f2) idx += 1;
    }
    *len = n + 1;
    return CP_OK;
}

CP_API float cp_fade_get(const cp_fade* f, int32_t frame) {
    if (!f || f->is_none) return 1.0f;
    int32_t i = frame - f->offset;
    int32_t len = (int32_t)f->len;
    if (i >= 0 && i < len) return f->values[i];
    if (!f->repeat) return i < 0 ? f->values[0] : f->values[len - 1];
    while (i < 0) i += len;
    i = i % len;
    return f->values[i];
}

CP_API int cp_outlier(cp_ctx* ctx, const cp_stack* st, const cp_frames* frames,
                      const cp_outlier_params* p, const cp_outputs* out, uint32_t flags) {
    if (!ctx || !st || !frames || !p || !out || !out->px)
        return fail(CP_ERR_INVALID_ARG, "null argument");
    if (st->ctx != ctx) return fail(CP_ERR_INVALID_ARG, "stack belongs to another context");
    OutlierPrep prep;
    int rc = prepare_outlier(st, frames, p, &prep);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    OutlierK k = prep.k;
    if ((rc = stage_outlier(ctx->arena, prep, prep.pos_off, s, &k))) return rc;
    struct { uint32_t out_w, out_h; } g{k.out_w, k.out_h};

    const uint64_t npx = (uint64_t)g.out_w * g.out_h;
    const uint64_t px_bytes = npx * st->channels * st->bps;
    const bool dev_out = flags & CP_FLAG_DEVICE_OUT;
    if (dev_out) {
        k.out_px = out->px;
        k.out_mask = out->mask;
        k.warnings = (unsigned long long*)out->warnings;
        k.stats = (PixelStats*)out->stats;
    } else {
        if ((rc = ctx->ensure(ctx->out_px, px_bytes))) return rc;
        k.out_px = ctx->out_px.ptr;
        if (out->mask) {
            if ((rc = ctx->ensure(ctx->out_mask, npx * st->channels))) return rc;
            k.out_mask = (uint8_t*)ctx->out_mask.ptr;
        }
        if (out->stats) {
            if ((rc = ctx->ensure(ctx->out_stats, npx * sizeof(cp_pixel_stats)))) return rc;
            k.stats = (PixelStats*)ctx->out_stats.ptr;
        }
        if ((rc = ctx->ensure(ctx->warn, 256))) return rc;
        k.warnings = (unsigned long long*)ctx->warn.ptr;
        CP_HIP_TRY(hipMemsetAsync(ctx->warn.ptr, 0, 8, s));
    }
    const uint64_t max_pos = *std::max_element(prep.pos_off.begin(), prep.pos_off.end());
    if ((rc = launch_outlier(ctx, st, k, max_pos, flags, s))) return rc;
    if (!dev_out) {
        CP_HIP_TRY(hipMemcpyAsync(out->px, k.out_px, px_bytes, hipMemcpyDeviceToHost, s));
        if (out->mask)
            CP_HIP_TRY(hipMemcpyAsync(out->mask, k.out_mask, npx * st->channels,
                                      hipMemcpyDeviceToHost, s));
        if (out->stats)
            CP_HIP_TRY(hipMemcpyAsync(out->stats, k.stats, npx * sizeof(cp_pixel_stats),
                                      hipMemcpyDeviceToHost, s));
        unsigned long long w = 0;
        CP_HIP_TRY(hipMemcpyAsync(&w, k.warnings, 8, hipMemcpyDeviceToHost, s));
        CP_HIP_TRY(hipStreamSynchronize(s));
        if (out->warnings) *out->warnings = w;
    }
    return CP_OK;
}

CP_API int cp_simple(cp_ctx* ctx, const cp_stack* st, const cp_frames* frames,
                     const float weights[4], const cp_fade* fade, int32_t darker, void* out_px,
                     uint32_t flags) {
    if (!ctx || !st || !frames || !weights || !out_px) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (st->ctx != ctx) return fail(CP_ERR_INVALID_ARG, "stack belongs to another context");
    int rc = check_weights(weights);
    if (rc) return rc;
    Geometry g;
    if ((rc = resolve_frames(st, frames, &g))) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    SimpleK k{};
    k.stack = st->dev;
    k.pos_off = (const uint64_t*)ctx->arena.put(g.pos_off.data(), 8 * g.n, s);
    if (!k.pos_off) return fail(CP_ERR_HIP, "parameter staging failed");
    k.n = g.n;
    k.stack_w = st->width;
    k.out_w = g.out_w;
    k.out_h = g.out_h;
    k.frame_offset = frames->frame_offset;
    k.darker = darker ? 1 : 0;
    for (int i = 0; i < 4; ++i) k.w[i] = weights[i];
    k.fade = fade_to_device(ctx, fade, &rc);
    if (rc) return rc;
    const uint64_t px_bytes = (uint64_t)g.out_w * g.out_h * st->channels * st->bps;
    const bool dev_out = flags & CP_FLAG_DEVICE_OUT;
    if (dev_out) {
        k.out_px = out_px;
    } else {
        if ((rc = ctx->ensure(ctx->out_px, px_bytes))) return rc;
        k.out_px = ctx->out_px.ptr;
    }
    if ((rc = launch_simple(ctx, st, k, s))) return rc;
    if (!dev_out) {
        CP_HIP_TRY(hipMemcpyAsync(out_px, k.out_px, px_bytes, hipMemcpyDeviceToHost, s));
        CP_HIP_TRY(hipStreamSynchronize(s));
    }
    return CP_OK;
}

CP_API int cp_crop_create(const int32_t* offsets, uint32_t n, uint32_t width, uint32_t height,
                          int32_t* crop_xy, uint32_t* out_w, uint32_t* out_h, int32_t* has_crop) {
    if (!offsets || !out_w || !out_h || !has_crop) return fail(CP_ERR_INVALID_ARG, "null argument");
    int32_t xmin = 0, ymin = 0, xmax = 0, ymax = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t x = offsets[2 * i], y = offsets[2 * i + 1];
        if (x < xmin) xmin = x;
        if (y < ymin) ymin = y;
        if (x > xmax) xmax = x;
        if (y > ymax) ymax = y;
    }
    if (xmin == 0 && ymin == 0 && xmax == 0 && ymax == 0) {
        *has_crop = 0;
        *out_w = width;
        *out_h = height;
        return CP_OK;
    }
    int32_t w = (int32_t)width + xmin - xmax, h = (int32_t)height + ymin - ymax;
    if (w <= 0 || h <= 0) return fail(CP_ERR_INVALID_ARG, "shake offsets leave an empty crop");
    if (!crop_xy) return fail(CP_ERR_INVALID_ARG, "null crop output");
    *has_crop = 1;
    *out_w = (uint32_t)w;
    *out_h = (uint32_t)h;
    for (uint32_t i = 0; i < n; ++i) {
        crop_xy[2 * i] = -xmin + offsets[2 * i];
        crop_xy[2 * i + 1] = -ymin + offsets[2 * i + 1];
    }
    return CP_OK;
}

CP_API int cp_video_windows(const cp_frame_range* vin, const cp_frame_range* vout,
                            uint32_t image_count, int32_t* win_start, int32_t* win_end,
                            int32_t* win_frame, int32_t cap, int32_t* n_out) {
    if (!vin || !vout || !n_out) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (vin->step == 0 || vout->step == 0) return fail(CP_ERR_INVALID_ARG, "frame range step must be > 0");
    int32_t v_lower;
    if (vout->has_start) v_lower = vout->start;
    else if (vin->has_start && vin->has_end) v_lower = -(vin->end - vin->start) + 1;
    else v_lower = 0;
    const int32_t v_upper = vout->has_end ? vout->end : (int32_t)image_count;
    int32_t nframes = (v_upper - v_lower) / (int32_t)vout->step;
    if (nframes < 0) nframes = 0;
    const int32_t step = (int32_t)vin->step;
    for (int32_t i = 0; i < nframes; ++i) {
        const int32_t frame = i * (int32_t)vout->step + v_lower;
        int32_t start = 0, end = (int32_t)image_count;
        if (vin->has_start) {
            int32_t st = frame + vin->start;
            while (st < 0) st += step;
            start = std::max(st % step, frame + vin->start);
        }
        if (vin->has_end)
            end = std::min((int32_t)image_count + (frame + vin->end) % step - step, frame + vin->end);
        if (i < cap && win_start && win_end && win_frame) {
            win_start[i] = start;
            win_end[i] = end;
            win_frame[i] = frame - v_lower;
        }
    }
    *n_out = nframes;
    return CP_OK;
}

}  // extern "C"
