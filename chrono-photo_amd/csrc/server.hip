// server.hip -- continuous-batching request loop (cp_server_*).
//
// A worker thread owns max_inflight stream slots and one upload stream.  Requests queue in
// arrival order; the worker admits the queue head whenever the server's HBM budget covers
// its device memory (uploaded frames and server-side outputs).  A request that brings host
// frames first uploads them on the upload stream; the worker watches that upload's event
// and only then hands the request to a slot, so no compute queue ever holds a wait on a
// copy (a slot stream waiting on an upload would also stall the other streams that share
// its hardware queue).  On its slot a request runs composite, download and a completion
// event; finished requests are retired by polling.  Requests run concurrently on different
// streams; each request's arithmetic is exactly that of a plain call, so results do not
// depend on what else is in flight.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "runtime.h"

using namespace cp;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    cp_request req{};
    std::vector<uint32_t> frames, sample;
    std::vector<int32_t> crop;
    std::vector<float> fade_o, fade_s;
    uint64_t need = 0;
    Clock::time_point t_submit, t_done;
    int status = 1;  // 1: pending, 0: done, < 0: error
    std::string error;
    cp_stack own_stack;
    hipEvent_t uploaded = nullptr;  // host frames on the device (requests without a stack)
    void* own_frames = nullptr;
    void* own_px = nullptr;
    void* own_mask = nullptr;
    uint64_t pixel_samples = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    ParamArena arena;
    Request* cur = nullptr;
};

// Copy by a compute queue between device memory and pinned, device-mapped host memory:
// 16-byte accesses when both ends allow them.
__global__ void __launch_bounds__(256) copy_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                   uint64_t bytes) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    const uint64_t i0 = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool wide = (((uintptr_t)src | (uintptr_t)dst) & 15u) == 0;
    const uint64_t n16 = wide ? bytes / 16 : 0;
    for (uint64_t i = i0; i < n16; i += stride)
        reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    for (uint64_t i = n16 * 16 + i0; i < bytes; i += stride) dst[i] = src[i];
}

// The device address of pinned host memory (hipHostMalloc), else null.
void* mapped_host(void* p) {
    hipPointerAttribute_t a;
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer || !a.hostPointer) return nullptr;
    return (uint8_t*)a.devicePointer + ((uint8_t*)p - (uint8_t*)a.hostPointer);
}

// Result download: through the DMA path, or, while uploads are in flight (a download
// queued behind a multi-megabyte upload in a DMA engine would hold its slot), by a copy
// kernel on the slot's own queue when the host buffer is pinned.
int download(void* host, const void* dev, uint64_t bytes, hipStream_t s, bool by_kernel, uint32_t* kernel_count) {
    void* d = by_kernel ? mapped_host(host) : nullptr;
    if (d) {
        const uint32_t blocks = (uint32_t)std::min<uint64_t>(1024, (bytes + 4095) / 4096 + 1);
        hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, (const uint8_t*)dev, (uint8_t*)d, bytes);
        ++*kernel_count;
        return hipGetLastError() == hipSuccess ? CP_OK : -1;
    }
    return hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s) == hipSuccess ? CP_OK : -1;
}

uint64_t request_bytes(const cp_request& r, uint32_t* out_w, uint32_t* out_h) {
    const cp_stack* st = r.job.stack;
    const uint32_t W = st ? st->width : r.width, H = st ? st->height : r.height;
    const uint32_t C = st ? st->channels : r.channels, B = st ? st->bps : r.bytes_per_sample;
    *out_w = r.job.frames.out_width ? r.job.frames.out_width : W;
    *out_h = r.job.frames.out_height ? r.job.frames.out_height : H;
    uint64_t bytes = 0;
    if (!st) bytes += (uint64_t)W * H * C * B * r.n_frames + kStackSlack;
    if (!r.job.out.px) bytes += (uint64_t)(*out_w) * (*out_h) * C * B;
    if (!r.job.out.mask && r.host_out_mask) bytes += (uint64_t)(*out_w) * (*out_h) * C;
    return bytes;
}

bool fade_ok(const cp_fade& f) { return f.is_none || (f.len > 0 && f.values != nullptr); }

}  // namespace

// Locking: `mu` guards what submitters and waiters share with the worker -- the arrival
// queue, the ticket map, each request's status and the published statistics.  Everything
// else (slots, uploads in transit, ready list, admitted bytes) belongs to the worker thread,
// which stages parameters, launches and polls events without holding `mu`.
struct cp_server {
    cp_ctx* ctx = nullptr;  // nullptr once the context was destroyed (an inert shell)
    uint64_t budget = 0;
    std::vector<Slot> slots;
    hipStream_t up_stream = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next = 1;
    bool stop = false, stopped = false;
    std::thread worker;
    cp_server_stats stats{};
    // worker-private
    std::deque<Request*> uploading, ready;  // admitted: frames in transit / waiting for a slot
    std::vector<std::pair<Request*, int>> finished;  // completed since the last locked phase
    uint64_t used = 0;
    cp_server_stats wstats{};               // the worker's counters, published under `mu`

    int upload(Request* r);
    int start(Request* r, Slot& slot);
    void release_slot(Slot& slot, int status);
    void release_request(Request* r, int status);
    void loop();
    void shutdown();
};

// Allocates a request's frames (with the slack the two-pixel tiles read past the last
// frame) and starts their upload on the upload stream.
int cp_server::upload(Request* r) {
    const cp_request& q = r->req;
    const uint64_t bytes = (uint64_t)q.width * q.height * q.channels * q.bytes_per_sample * q.n_frames;
    if (hipMallocAsync(&r->own_frames, bytes + kStackSlack, up_stream) != hipSuccess)
        return fail(CP_ERR_HIP, "server: frame allocation of %llu bytes failed", (unsigned long long)bytes);
    if (hipMemcpyAsync(r->own_frames, q.host_frames, bytes, hipMemcpyHostToDevice, up_stream) != hipSuccess)
        return fail(CP_ERR_HIP, "server: frame upload failed");
    if (hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(r->uploaded, up_stream) != hipSuccess)
        return fail(CP_ERR_HIP, "server: upload event failed");
    return CP_OK;
}

int cp_server::start(Request* r, Slot& slot) {
    hipStream_t s = slot.stream;
    cp_request& q = r->req;
    cp_job& j = q.job;
    const cp_stack* st = j.stack;
    if (!st) {  // frames uploaded by upload(); the worker saw its event complete
        cp_stack& o = r->own_stack;
        o.ctx = ctx;
        o.width = q.width;
        o.height = q.height;
        o.channels = q.channels;
        o.bps = q.bytes_per_sample;
        o.capacity = o.count = q.n_frames;
        o.dev = (uint8_t*)r->own_frames;
        o.alloc_bytes = o.frame_bytes() * q.n_frames + kStackSlack;
        o.owned = false;
        st = &o;
    }
    uint32_t ow = 0, oh = 0;
    (void)request_bytes(q, &ow, &oh);
    const uint64_t px_bytes = (uint64_t)ow * oh * st->channels * st->bps;
    const uint64_t mask_bytes = (uint64_t)ow * oh * st->channels;
    void* px = j.out.px;
    uint8_t* mask = j.out.mask;
    if (!px) {
        if (hipMallocAsync(&r->own_px, px_bytes, s) != hipSuccess) return fail(CP_ERR_HIP, "server: output allocation failed");
        px = r->own_px;
    }
    if (!mask && q.host_out_mask) {
        if (hipMallocAsync(&r->own_mask, mask_bytes, s) != hipSuccess) return fail(CP_ERR_HIP, "server: mask allocation failed");
        mask = (uint8_t*)r->own_mask;
    }
    int rc;
    if (j.mode == CP_MODE_OUTLIER) {
        OutlierPrep prep;
        if ((rc = prepare_outlier(st, &j.frames, &j.outlier, &prep))) return rc;
        OutlierK k = prep.k;
        if ((rc = stage_outlier(slot.arena, prep, prep.pos_off, s, &k))) return rc;
        k.out_px = px;
        k.out_mask = mask;
        k.warnings = (unsigned long long*)j.out.warnings;
        k.stats = (PixelStats*)j.out.stats;
        const uint64_t max_pos = *std::max_element(prep.pos_off.begin(), prep.pos_off.end());
        if ((rc = launch_outlier(ctx, st, k, max_pos, 0, s))) return rc;
        r->pixel_samples = (uint64_t)k.n * k.out_w * k.out_h;
    } else {
        if ((rc = check_weights(j.weights))) return rc;
        Geometry g;
        if ((rc = resolve_frames(st, &j.frames, &g))) return rc;
        SimpleK k{};
        k.stack = st->dev;
        k.pos_off = (const uint64_t*)slot.arena.put(g.pos_off.data(), 8 * g.n, s);
        if (!k.pos_off) return fail(CP_ERR_HIP, "server: parameter staging failed");
        k.n = g.n;
        k.stack_w = st->width;
        k.out_w = g.out_w;
        k.out_h = g.out_h;
        k.frame_offset = j.frames.frame_offset;
        k.darker = j.mode == CP_MODE_DARKER;
        for (int i = 0; i < 4; ++i) k.w[i] = j.weights[i];
        k.fade.is_none = j.fade.is_none ? 1 : 0;
        if (!j.fade.is_none) {  // cp_server_submit checked the table
            k.fade.repeat = j.fade.repeat;
            k.fade.absolute = j.fade.absolute;
            k.fade.offset = j.fade.offset;
            k.fade.len = j.fade.len;
            k.fade.values = (const float*)slot.arena.put(j.fade.values, 4 * j.fade.len, s);
            if (!k.fade.values) return fail(CP_ERR_HIP, "server: parameter staging failed");
        }
        k.out_px = px;
        if ((rc = launch_simple(ctx, st, k, s))) return rc;
        r->pixel_samples = (uint64_t)k.n * k.out_w * k.out_h;
    }
    const bool by_kernel = !uploading.empty();
    if (q.host_out_px && download(q.host_out_px, px, px_bytes, s, by_kernel, &wstats.kernel_downloads) != CP_OK)
        return fail(CP_ERR_HIP, "server: result download failed");
    if (q.host_out_mask && mask &&
        download(q.host_out_mask, mask, mask_bytes, s, by_kernel, &wstats.kernel_downloads) != CP_OK)
        return fail(CP_ERR_HIP, "server: mask download failed");
    if (hipEventRecord(slot.done, s) != hipSuccess) return fail(CP_ERR_HIP, "server: event record failed");
    return CP_OK;
}

// A slot's request is done (or failed): free its device memory in stream order.
void cp_server::release_slot(Slot& slot, int status) {
    Request* r = slot.cur;
    hipStream_t s = slot.stream;
    if (r->own_frames) (void)hipFreeAsync(r->own_frames, s);
    if (r->own_px) (void)hipFreeAsync(r->own_px, s);
    if (r->own_mask) (void)hipFreeAsync(r->own_mask, s);
    r->own_frames = r->own_px = r->own_mask = nullptr;
    slot.cur = nullptr;
    release_request(r, status);
}

// Releases what is left of a request (not on a slot) and queues its completion.
void cp_server::release_request(Request* r, int status) {
    if (r->own_frames) {  // an upload that failed or never reached a slot
        (void)hipFreeAsync(r->own_frames, up_stream);
        (void)hipStreamSynchronize(up_stream);
        r->own_frames = nullptr;
    }
    if (r->uploaded) {
        (void)hipEventDestroy(r->uploaded);
        r->uploaded = nullptr;
    }
    used -= r->need;
    r->t_done = Clock::now();
    if (status == CP_OK) {
        wstats.completed++;
        wstats.pixel_samples += r->pixel_samples;
    }
    finished.emplace_back(r, status);
}

void cp_server::loop() {
    (void)hipSetDevice(ctx->device);
    std::vector<Request*> admitted;
    for (;;) {
        // ---- device progress, without the lock
        bool busy = false;
        for (auto& sl : slots) {
            if (!sl.cur) continue;
            const hipError_t e = hipEventQuery(sl.done);
            if (e == hipSuccess) {
                release_slot(sl, CP_OK);
            } else if (e != hipErrorNotReady) {
                sl.cur->error = hipGetErrorString(e);
                release_slot(sl, CP_ERR_HIP);
            } else {
                busy = true;
            }
        }
        for (auto it = uploading.begin(); it != uploading.end();) {  // arrived frames become ready
            Request* r = *it;
            const hipError_t e = hipEventQuery(r->uploaded);
            if (e == hipSuccess) {
                ready.push_back(r);
                it = uploading.erase(it);
            } else if (e != hipErrorNotReady) {
                r->error = hipGetErrorString(e);
                it = uploading.erase(it);
                release_request(r, CP_ERR_HIP);
            } else {
                busy = true;
                ++it;
            }
        }
        // ---- shared state: completions, statistics, admission in arrival order
        admitted.clear();
        {
            std::unique_lock<std::mutex> lock(mu);
            for (auto& f : finished) f.first->status = f.second;
            bool any_done = !finished.empty();
            finished.clear();
            while (!queue.empty()) {
                Request* r = queue.front();
                if (budget && r->need > budget) {
                    queue.pop_front();
                    r->status = CP_ERR_OUT_OF_HBM_BUDGET;
                    r->error = "request needs more device memory than the server budget";
                    r->t_done = Clock::now();
                    wstats.rejected++;
                    any_done = true;
                    continue;
                }
                if (budget && used + r->need > budget) break;  // wait for memory to free up
                // at most one upload in transit per slot: without a budget this bounds the memory
                if (!r->req.job.stack && uploading.size() + admitted.size() >= slots.size()) break;
                queue.pop_front();
                used += r->need;
                if (used > wstats.peak_bytes) wstats.peak_bytes = used;
                admitted.push_back(r);
            }
            const uint64_t submitted = stats.submitted;
            stats = wstats;
            stats.submitted = submitted;
            if (any_done) cv_done.notify_all();
            const bool idle = !busy && admitted.empty() && ready.empty() && uploading.empty();
            if (idle) {
                if (stop && queue.empty()) break;
                if (queue.empty()) {
                    cv_work.wait(lock, [&] { return stop || !queue.empty(); });
                    continue;
                }
            }
        }
        // ---- staging and launches, without the lock
        for (Request* r : admitted) {
            if (r->req.job.stack) {
                ready.push_back(r);
                continue;
            }
            const int rc = upload(r);
            if (rc != CP_OK) {
                r->error = cp_last_error();
                release_request(r, rc);
                continue;
            }
            uploading.push_back(r);
        }
        while (!ready.empty()) {
            Slot* free_slot = nullptr;
            for (auto& sl : slots)
                if (!sl.cur) { free_slot = &sl; break; }
            if (!free_slot) break;
            Request* r = ready.front();
            ready.pop_front();
            free_slot->cur = r;
            const int rc = start(r, *free_slot);
            if (rc != CP_OK) {
                r->error = cp_last_error();
                (void)hipStreamSynchronize(free_slot->stream);
                release_slot(*free_slot, rc);
                continue;
            }
            uint32_t inflight = 0;
            for (auto& sl : slots) inflight += sl.cur != nullptr;
            if (inflight > wstats.peak_inflight) wstats.peak_inflight = inflight;
        }
        if (finished.empty()) {  // nothing to publish: give the device a moment (a submit wakes us)
            std::unique_lock<std::mutex> lock(mu);
            if (!stop) cv_work.wait_for(lock, std::chrono::microseconds(20));
        }
    }
}

// Drains the queue and in-flight work, stops the worker and frees the streams.
void cp_server::shutdown() {
    if (stopped) return;
    {
        std::lock_guard<std::mutex> lock(mu);
        stop = true;
        cv_work.notify_all();
    }
    if (worker.joinable()) worker.join();
    for (auto& sl : slots) {
        (void)hipStreamSynchronize(sl.stream);
        (void)hipStreamDestroy(sl.stream);
        (void)hipEventDestroy(sl.done);
        sl.arena.release();
    }
    (void)hipStreamSynchronize(up_stream);
    (void)hipStreamDestroy(up_stream);
    stopped = true;
}

void cp::server_orphan(cp_server* v) {
    v->shutdown();
    v->ctx = nullptr;
}

extern "C" CP_API int cp_server_create(cp_ctx* ctx, uint64_t hbm_budget_bytes, uint32_t max_inflight,
                                       cp_server** out) {
    if (!ctx || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (max_inflight == 0) max_inflight = 4;
    if (max_inflight > 16) return fail(CP_ERR_INVALID_ARG, "at most 16 in-flight slots");
    auto srv = std::make_unique<cp_server>();
    srv->ctx = ctx;
    srv->budget = hbm_budget_bytes;
    srv->slots.resize(max_inflight);
    (void)hipSetDevice(ctx->device);
    if (hipStreamCreateWithFlags(&srv->up_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(CP_ERR_HIP, "server: upload stream creation failed");
    for (auto& sl : srv->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess || sl.arena.init(1u << 20) != 0) {
            for (auto& t : srv->slots) {
                if (t.stream) (void)hipStreamDestroy(t.stream);
                if (t.done) (void)hipEventDestroy(t.done);
                t.arena.release();
            }
            (void)hipStreamDestroy(srv->up_stream);
            return fail(CP_ERR_HIP, "server: stream slot creation failed");
        }
    }
    cp_server* s = srv.release();
    s->worker = std::thread([s] { s->loop(); });
    ctx->attach(s);
    *out = s;
    return CP_OK;
}

extern "C" CP_API int cp_server_submit(cp_server* srv, const cp_request* req, uint64_t* ticket) {
    if (!srv || !req || !ticket) return fail(CP_ERR_INVALID_ARG, "null argument");
    if (!srv->ctx) return fail(CP_ERR_INVALID_ARG, "the server's context was destroyed");
    const cp_job& j = req->job;
    if (j.mode < CP_MODE_OUTLIER || j.mode > CP_MODE_DARKER) return fail(CP_ERR_INVALID_ARG, "bad mode");
    // fade tables are read on the device: an empty or missing table is refused here, before
    // anything is queued (cp_simple and cp_outlier refuse it the same way)
    if (j.mode == CP_MODE_OUTLIER ? !fade_ok(j.outlier.fade) : !fade_ok(j.fade))
        return fail(CP_ERR_INVALID_ARG, "fade table is empty");
    if (!j.stack) {
        if (!req->host_frames || req->n_frames == 0 || req->width == 0 || req->height == 0)
            return fail(CP_ERR_INVALID_ARG, "request without stack needs host frames");
        if (req->channels < 1 || req->channels > 4 || (req->bytes_per_sample != 1 && req->bytes_per_sample != 2))
            return fail(CP_ERR_INVALID_ARG, "bad frame layout");
    } else if (j.stack->ctx != srv->ctx) {
        return fail(CP_ERR_INVALID_ARG, "stack belongs to another context");
    }
    if (!j.out.px && !req->host_out_px) return fail(CP_ERR_INVALID_ARG, "request has no output");
    auto r = std::make_unique<Request>();
    r->req = *req;
    const uint32_t n = j.frames.n;
    cp_job& jj = r->req.job;
    if (j.frames.frames) { r->frames.assign(j.frames.frames, j.frames.frames + n); jj.frames.frames = r->frames.data(); }
    if (j.frames.crop_xy) { r->crop.assign(j.frames.crop_xy, j.frames.crop_xy + 2 * n); jj.frames.crop_xy = r->crop.data(); }
    if (j.outlier.sample_pos) {
        r->sample.assign(j.outlier.sample_pos, j.outlier.sample_pos + j.outlier.n_sample);
        jj.outlier.sample_pos = r->sample.data();
    }
    if (!j.outlier.fade.is_none && j.outlier.fade.values) {
        r->fade_o.assign(j.outlier.fade.values, j.outlier.fade.values + j.outlier.fade.len);
        jj.outlier.fade.values = r->fade_o.data();
    }
    if (!j.fade.is_none && j.fade.values) {
        r->fade_s.assign(j.fade.values, j.fade.values + j.fade.len);
        jj.fade.values = r->fade_s.data();
    }
    uint32_t ow, oh;
    r->need = request_bytes(r->req, &ow, &oh);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lock(srv->mu);
    r->ticket = srv->next++;
    *ticket = r->ticket;
    srv->queue.push_back(r.get());
    srv->all[r->ticket] = std::move(r);
    srv->stats.submitted++;
    srv->cv_work.notify_one();
    return CP_OK;
}

extern "C" CP_API int cp_server_poll(cp_server* srv, uint64_t ticket, int32_t* done) {
    if (!srv || !done) return fail(CP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lock(srv->mu);
    auto it = srv->all.find(ticket);
    if (it == srv->all.end()) return fail(CP_ERR_INVALID_ARG, "unknown ticket");
    *done = it->second->status != 1;
    return CP_OK;
}

extern "C" CP_API int cp_server_wait(cp_server* srv, uint64_t ticket, double* latency_ms) {
    if (!srv) return fail(CP_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lock(srv->mu);
    auto it = srv->all.find(ticket);
    if (it == srv->all.end()) return fail(CP_ERR_INVALID_ARG, "unknown ticket");
    Request* r = it->second.get();
    srv->cv_done.wait(lock, [&] { return r->status != 1; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int status = r->status;
    std::string err = r->error;
    srv->all.erase(it);
    if (status != CP_OK) return fail(status, "request %llu: %s", (unsigned long long)ticket, err.c_str());
    return CP_OK;
}

extern "C" CP_API int cp_server_stats_get(cp_server* srv, cp_server_stats* out) {
    if (!srv || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lock(srv->mu);
    *out = srv->stats;
    return CP_OK;
}

extern "C" CP_API int cp_server_destroy(cp_server* srv) {
    if (!srv) return CP_OK;
    if (srv->ctx) {  // else its context was destroyed first and already shut it down
        srv->ctx->detach(srv);
        srv->shutdown();
    }
    delete srv;
    return CP_OK;
}
