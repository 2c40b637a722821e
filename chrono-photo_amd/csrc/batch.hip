// batch.hip -- cp_batch_run and the serving loop (placeholder: sequential launches).
#include "runtime.h"

using namespace cp;

extern "C" CP_API int cp_batch_run(cp_ctx* ctx, const cp_job* jobs, uint32_t n_jobs, uint32_t flags) {
    if (!ctx || (!jobs && n_jobs)) return fail(CP_ERR_INVALID_ARG, "null argument");
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const cp_job& j = jobs[i];
        int rc;
        if (j.mode == CP_MODE_OUTLIER)
            rc = cp_outlier(ctx, j.stack, &j.frames, &j.outlier, &j.out, flags | CP_FLAG_DEVICE_OUT);
        else
            rc = cp_simple(ctx, j.stack, &j.frames, j.weights, &j.fade, j.mode == CP_MODE_DARKER,
                           j.out.px, flags | CP_FLAG_DEVICE_OUT);
        if (rc) return rc;
    }
    return CP_OK;
}

// ---------------------------------------------------------------------------- serving
#include <chrono>
#include <unordered_map>

struct cp_server {
    cp_ctx* ctx = nullptr;
    uint64_t budget = 0;
    uint32_t max_inflight = 0;
    uint64_t next = 1;
    struct Ticket {
        hipEvent_t done;
        std::chrono::steady_clock::time_point t0;
    };
    std::unordered_map<uint64_t, Ticket> tickets;
};

extern "C" CP_API int cp_server_create(cp_ctx* ctx, uint64_t hbm_budget_bytes, uint32_t max_inflight,
                                       cp_server** out) {
    if (!ctx || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    cp_server* s = new cp_server();
    s->ctx = ctx;
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight ? max_inflight : 8;
    *out = s;
    return CP_OK;
}

extern "C" CP_API int cp_server_submit(cp_server* srv, const cp_job* job, uint64_t* ticket) {
    if (!srv || !job || !ticket) return fail(CP_ERR_INVALID_ARG, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    int rc = cp_batch_run(srv->ctx, job, 1, CP_FLAG_DEVICE_OUT);
    if (rc) return rc;
    cp_server::Ticket t;
    t.t0 = t0;
    CP_HIP_TRY(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
    CP_HIP_TRY(hipEventRecord(t.done, srv->ctx->stream));
    *ticket = srv->next++;
    srv->tickets[*ticket] = t;
    return CP_OK;
}

extern "C" CP_API int cp_server_poll(cp_server* srv, uint64_t ticket, int32_t* done) {
    if (!srv || !done) return fail(CP_ERR_INVALID_ARG, "null argument");
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(CP_ERR_INVALID_ARG, "unknown ticket");
    hipError_t e = hipEventQuery(it->second.done);
    *done = e == hipSuccess;
    if (e != hipSuccess && e != hipErrorNotReady) return fail(CP_ERR_HIP, "%s", hipGetErrorString(e));
    return CP_OK;
}

extern "C" CP_API int cp_server_wait(cp_server* srv, uint64_t ticket, double* latency_ms) {
    if (!srv) return fail(CP_ERR_INVALID_ARG, "null argument");
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(CP_ERR_INVALID_ARG, "unknown ticket");
    CP_HIP_TRY(hipEventSynchronize(it->second.done));
    auto t1 = std::chrono::steady_clock::now();
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(t1 - it->second.t0).count();
    (void)hipEventDestroy(it->second.done);
    srv->tickets.erase(it);
    return CP_OK;
}

extern "C" CP_API int cp_server_destroy(cp_server* srv) {
    if (!srv) return CP_OK;
    for (auto& kv : srv->tickets) {
        (void)hipEventSynchronize(kv.second.done);
        (void)hipEventDestroy(kv.second.done);
    }
    delete srv;
    return CP_OK;
}
