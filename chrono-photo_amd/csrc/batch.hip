// batch.hip -- batched launches (cp_batch_*) and the kernel autotuner.
//
// cp_batch_create stages every job's per-call tables (frame offsets, sample mask and offsets,
// fade tables) once in one device allocation owned by the batch, and plans each outlier job
// as its own launch would run it (plan_outlier: the tuned kernel and shape).  The jobs whose
// plan has a group form (k_batch.h: every packed-byte and 16-bit-halves shape) are grouped by
// kernel; each group is one launch over all its jobs' tiles.  The rest (the generic kernel's
// jobs, darker/lighter) are launched one by one.  The launches are spread, biggest first,
// over the context stream and three side streams forked from it and joined back; the whole
// sequence is captured once into a HIP graph, so a replay is one graph launch whose branches
// the device runs side by side.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "k_batch.h"
#include "kernels_fast.h"
#include "kernels_u8.h"
#include "runtime.h"

using namespace cp;

namespace {

constexpr int kSideStreams = 3;

// Device-table offsets of one job's per-call tables.
struct TFix {
    uint64_t pos = 0, fade = 0, mask = 0, samp = 0;
    bool has_fade = false, has_mask = false;
};

// A job launched on its own, with kernel arguments pointing into the batch's tables.
struct DirectJob {
    bool outlier = true;
    const cp_stack* st = nullptr;
    OutlierK ko{};
    uint64_t max_pos = 0;
    SimpleK ks{};
    TFix fix;
    uint64_t cost = 0;
    int stream = 0;  // 0: the context stream, 1..3: side streams
};

// The jobs of one group launch: one kernel (family, sample width, channels, shape, threshold
// kind).
struct PGroup {
    const void* kernel = nullptr;
    int family = 0, shape = -1;
    uint32_t bps = 0, channels = 0;
    bool absolute = false;
    std::vector<BatchJobK> jobs;
    std::vector<TFix> fix;
    std::vector<uint64_t> cost;
    uint64_t work = 0;
    uint32_t total = 0;  // tiles of all its jobs: the launch's grid
    uint64_t jobs_off = 0;
    int stream = 0;
};

// Host image of the batch's device tables: blobs appended at 256-byte aligned offsets.
struct Tables {
    std::vector<uint8_t> host;
    uint64_t put(const void* data, uint64_t bytes) {
        const uint64_t off = (host.size() + 255) & ~uint64_t(255);
        host.resize(off + std::max<uint64_t>(bytes, 1), 0);
        if (data && bytes) memcpy(host.data() + off, data, bytes);
        return off;
    }
};

void apply_fix(OutlierK& k, const TFix& f, uint8_t* base) {
    k.pos_off = (const uint64_t*)(base + f.pos);
    if (f.has_fade) k.fade.values = (const float*)(base + f.fade);
    if (f.has_mask) {
        k.sample_mask = (const uint32_t*)(base + f.mask);
        k.samp_off = (const uint64_t*)(base + f.samp);
    }
}

}  // namespace

struct cp_batch {
    cp_ctx* ctx = nullptr;
    std::vector<PGroup> pgroups;
    std::vector<DirectJob> direct;
    void* d_tables = nullptr;
    uint64_t table_bytes = 0;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork = nullptr, join[kSideStreams] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Device resources of a batch (graph, tables, streams); the context stream is drained first.
static void batch_release(cp_batch* b) {
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto& st : b->side)
        if (st) (void)hipStreamSynchronize(st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_tables) {
        (void)hipFree(b->d_tables);
        b->ctx->uncharge(b->table_bytes);
    }
    for (auto& st : b->side)
        if (st) (void)hipStreamDestroy(st);
    for (auto& e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    b->exec = nullptr;
    b->graph = nullptr;
    b->d_tables = nullptr;
    b->fork = nullptr;
    for (int i = 0; i < kSideStreams; ++i) b->side[i] = nullptr, b->join[i] = nullptr;
    b->pgroups.clear();
    b->direct.clear();
}

void cp::batch_orphan(cp_batch* b) {
    batch_release(b);
    b->ctx = nullptr;
}

// The launch sequence: fork the side streams from the context stream, each launch on its
// stream, join.
static int batch_enqueue(cp_batch* b) {
    cp_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    uint8_t* base = (uint8_t*)b->d_tables;
    if (hipEventRecord(b->fork, s) != hipSuccess) return fail(CP_ERR_HIP, "batch: fork event failed");
    for (auto& st : b->side)
        if (hipStreamWaitEvent(st, b->fork, 0) != hipSuccess) return fail(CP_ERR_HIP, "batch: fork wait failed");
    auto stream = [&](int i) { return i == 0 ? s : b->side[i - 1]; };
    for (auto& g : b->pgroups) {
        const BatchJobK* jobs = (const BatchJobK*)(base + g.jobs_off);
        uint32_t njobs = (uint32_t)g.jobs.size();
        void* args[] = {&jobs, &njobs};
        if (hipLaunchKernel(g.kernel, dim3(g.total), dim3(256), args, 0, stream(g.stream)) != hipSuccess)
            return fail(CP_ERR_HIP, "batch: group launch: %s", hipGetErrorString(hipGetLastError()));
    }
    for (const DirectJob& d : b->direct) {
        hipStream_t ss = stream(d.stream);
        const int rc = d.outlier ? launch_outlier(ctx, d.st, d.ko, d.max_pos, 0, ss) : launch_simple(ctx, d.st, d.ks, ss);
        if (rc) return rc;
    }
    for (int i = 0; i < kSideStreams; ++i) {
        if (hipEventRecord(b->join[i], b->side[i]) != hipSuccess || hipStreamWaitEvent(s, b->join[i], 0) != hipSuccess)
            return fail(CP_ERR_HIP, "batch: join failed");
    }
    return CP_OK;
}

extern "C" CP_API int cp_batch_destroy(cp_batch* b) {
    if (!b) return CP_OK;
    if (b->ctx) {
        b->ctx->detach(b);
        batch_release(b);
    }
    delete b;
    return CP_OK;
}

extern "C" CP_API int cp_batch_create(cp_ctx* ctx, const cp_job* jobs, uint32_t n_jobs, uint32_t flags,
                                      cp_batch** out) {
    if (!ctx || !out || (!jobs && n_jobs)) return fail(CP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    // a batch that fails half-way releases what it already holds
    struct Release {
        void operator()(cp_batch* p) const {
            if (p->ctx) batch_release(p);
            delete p;
        }
    };
    std::unique_ptr<cp_batch, Release> b(new cp_batch());
    b->ctx = ctx;
    Tables tab;
    std::vector<PGroup> groups;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const cp_job& j = jobs[i];
        if (!j.stack || !j.out.px) return fail(CP_ERR_INVALID_ARG, "job %u: null stack or output", i);
        if (j.stack->ctx != ctx) return fail(CP_ERR_INVALID_ARG, "job %u: stack belongs to another context", i);
        if (j.mode < CP_MODE_OUTLIER || j.mode > CP_MODE_DARKER) return fail(CP_ERR_INVALID_ARG, "job %u: bad mode", i);
        const cp_stack* st = j.stack;
        if (j.mode == CP_MODE_OUTLIER) {
            OutlierPrep prep;
            int rc = prepare_outlier(st, &j.frames, &j.outlier, &prep);
            if (rc) return rc;
            OutlierK k = prep.k;
            k.out_px = j.out.px;
            k.out_mask = j.out.mask;
            k.warnings = (unsigned long long*)j.out.warnings;
            k.stats = (PixelStats*)j.out.stats;
            const uint64_t max_pos = *std::max_element(prep.pos_off.begin(), prep.pos_off.end());
            TFix f;
            f.pos = tab.put(prep.pos_off.data(), 8 * prep.pos_off.size());
            f.has_fade = !prep.fade_values.empty();
            if (f.has_fade) f.fade = tab.put(prep.fade_values.data(), 4 * prep.fade_values.size());
            f.has_mask = !prep.sample_mask.empty();
            if (f.has_mask) {
                f.mask = tab.put(prep.sample_mask.data(), 4 * prep.sample_mask.size());
                std::vector<uint64_t> so(prep.sample_pos.size());
                for (size_t q = 0; q < so.size(); ++q) so[q] = prep.pos_off[prep.sample_pos[q]];
                f.samp = tab.put(so.data(), 8 * so.size());
                // the plan only asks whether a sample subset is there
                k.sample_mask = (const uint32_t*)uintptr_t(1);
                k.samp_off = (const uint64_t*)uintptr_t(1);
            }
            const uint64_t cost = (uint64_t)k.n * k.out_w * k.out_h * st->channels * st->bps;
            const OutlierPlan plan = plan_outlier(ctx, st, k, max_pos, 0);
            const bool absolute = k.th.absolute != 0;
            const void* kern = plan.family == kPlanGeneric
                                   ? nullptr
                                   : group_kernel(plan.family, plan.shape, st->bps, st->channels, absolute);
            if (kern) {
                PGroup* g = nullptr;
                for (auto& q : groups)
                    if (q.kernel == kern) g = &q;
                if (!g) {
                    groups.emplace_back();
                    g = &groups.back();
                    g->kernel = kern;
                    g->family = plan.family;
                    g->shape = plan.shape;
                    g->bps = st->bps;
                    g->channels = st->channels;
                    g->absolute = absolute;
                }
                BatchJobK bj{};
                bj.k = k;
                const uint32_t P = group_tile_pixels(plan.family, plan.shape);
                bj.tiles_x = (k.out_w + P - 1) / P;
                bj.tiles = bj.tiles_x * k.out_h;
                g->jobs.push_back(bj);
                g->fix.push_back(f);
                g->cost.push_back(cost);
                g->work += cost;
                continue;
            }
            DirectJob d;
            d.outlier = true;
            d.st = st;
            d.ko = k;
            d.max_pos = max_pos;
            d.fix = f;
            d.cost = cost;
            b->direct.push_back(d);
            continue;
        }
        int rc = check_weights(j.weights);
        if (rc) return rc;
        if (!j.fade.is_none && (j.fade.len == 0 || !j.fade.values))
            return fail(CP_ERR_INVALID_ARG, "job %u: fade table is empty", i);
        Geometry geo;
        if ((rc = resolve_frames(st, &j.frames, &geo))) return rc;
        DirectJob d;
        d.outlier = false;
        d.st = st;
        SimpleK& k = d.ks;
        k.stack = st->dev;
        k.n = geo.n;
        k.stack_w = st->width;
        k.out_w = geo.out_w;
        k.out_h = geo.out_h;
        k.frame_offset = j.frames.frame_offset;
        k.darker = j.mode == CP_MODE_DARKER;
        for (int c = 0; c < 4; ++c) k.w[c] = j.weights[c];
        k.fade.is_none = j.fade.is_none ? 1 : 0;
        k.fade.repeat = j.fade.repeat;
        k.fade.absolute = j.fade.absolute;
        k.fade.offset = j.fade.offset;
        k.fade.len = j.fade.len;
        k.out_px = j.out.px;
        d.fix.pos = tab.put(geo.pos_off.data(), 8 * geo.pos_off.size());
        d.fix.has_fade = !j.fade.is_none;
        if (d.fix.has_fade) d.fix.fade = tab.put(j.fade.values, 4 * j.fade.len);
        d.cost = (uint64_t)geo.n * geo.out_w * geo.out_h * st->channels * st->bps;
        b->direct.push_back(d);
    }
    // groups: biggest jobs first, then their tiles
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        PGroup& g = groups[gi];
        std::vector<size_t> ord(g.jobs.size());
        for (size_t q = 0; q < ord.size(); ++q) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) { return g.cost[a] > g.cost[c]; });
        std::vector<BatchJobK> sorted;
        std::vector<TFix> sfix;
        uint64_t tiles = 0;
        for (size_t q : ord) {
            sorted.push_back(g.jobs[q]);
            sfix.push_back(g.fix[q]);
            sorted.back().tile0 = (uint32_t)tiles;
            tiles += sorted.back().tiles;
        }
        if (tiles >= 0x7FFFFFFFull) return fail(CP_ERR_CAPACITY, "batch has too many tiles");
        g.jobs.swap(sorted);
        g.fix.swap(sfix);
        g.total = (uint32_t)tiles;
        g.jobs_off = tab.put(nullptr, sizeof(BatchJobK) * g.jobs.size());
    }
    // streams: launches by decreasing work, each onto the least-loaded stream
    {
        std::vector<std::pair<uint64_t, int>> order;  // (work, index: groups first, then direct jobs)
        for (size_t q = 0; q < groups.size(); ++q) order.push_back({groups[q].work, (int)q});
        for (size_t q = 0; q < b->direct.size(); ++q) order.push_back({b->direct[q].cost, (int)(groups.size() + q)});
        std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& c) { return a.first > c.first; });
        uint64_t load[1 + kSideStreams] = {};
        for (const auto& o : order) {
            int best = 0;
            for (int q = 1; q <= kSideStreams; ++q)
                if (load[q] < load[best]) best = q;
            load[best] += o.first;
            if (o.second < (int)groups.size()) groups[o.second].stream = best;
            else b->direct[o.second - groups.size()].stream = best;
        }
    }
    b->pgroups.swap(groups);
    // device tables
    const uint64_t bytes = tab.host.size();
    {
        int rc = ctx->charge(bytes);
        if (rc) return rc;
        (void)hipSetDevice(ctx->device);
        if (hipMalloc(&b->d_tables, bytes) != hipSuccess) {
            ctx->uncharge(bytes);
            return fail(CP_ERR_HIP, "hipMalloc of %llu batch table bytes failed", (unsigned long long)bytes);
        }
        b->table_bytes = bytes;
        uint8_t* dbase = (uint8_t*)b->d_tables;
        for (DirectJob& d : b->direct) {
            if (d.outlier) {
                apply_fix(d.ko, d.fix, dbase);
            } else {
                d.ks.pos_off = (const uint64_t*)(dbase + d.fix.pos);
                if (d.fix.has_fade) d.ks.fade.values = (const float*)(dbase + d.fix.fade);
            }
        }
        for (PGroup& g : b->pgroups) {
            for (size_t q = 0; q < g.jobs.size(); ++q) apply_fix(g.jobs[q].k, g.fix[q], dbase);
            memcpy(tab.host.data() + g.jobs_off, g.jobs.data(), sizeof(BatchJobK) * g.jobs.size());
        }
        CP_HIP_TRY(hipMemcpy(b->d_tables, tab.host.data(), bytes, hipMemcpyHostToDevice));
    }
    (void)hipSetDevice(ctx->device);
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return fail(CP_ERR_HIP, "batch: event creation failed");
    for (int q = 0; q < kSideStreams; ++q)
        if (hipStreamCreateWithFlags(&b->side[q], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->join[q], hipEventDisableTiming) != hipSuccess)
            return fail(CP_ERR_HIP, "batch: side stream creation failed");
    if (!(flags & CP_FLAG_NO_GRAPH) && (!b->pgroups.empty() || !b->direct.empty())) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        CP_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int lrc = batch_enqueue(b.get());
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (lrc || e != hipSuccess || !graph)
            return fail(CP_ERR_HIP, "batch graph capture failed: %s", hipGetErrorString(e));
        b->graph = graph;
        CP_HIP_TRY(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
    }
    ctx->attach(b.get());
    *out = b.release();
    return CP_OK;
}

extern "C" CP_API int cp_batch_launch(cp_batch* b) {
    if (!b) return fail(CP_ERR_INVALID_ARG, "null batch");
    if (!b->ctx) return fail(CP_ERR_INVALID_ARG, "the batch's context was destroyed");
    cp_ctx* ctx = b->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (b->exec) {
        CP_HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
        return CP_OK;
    }
    return batch_enqueue(b);
}

extern "C" CP_API int cp_batch_info_get(const cp_batch* b, cp_batch_info* out) {
    if (!b || !out) return fail(CP_ERR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    out->group_launches = (uint32_t)b->pgroups.size();
    for (const PGroup& g : b->pgroups) out->group_jobs += (uint32_t)g.jobs.size();
    out->direct_jobs = (uint32_t)b->direct.size();
    return CP_OK;
}

extern "C" CP_API int cp_batch_run(cp_ctx* ctx, const cp_job* jobs, uint32_t n_jobs, uint32_t flags) {
    cp_batch* b = nullptr;
    int rc = cp_batch_create(ctx, jobs, n_jobs, flags | CP_FLAG_NO_GRAPH, &b);
    if (rc) return rc;
    rc = cp_batch_launch(b);
    if (rc == CP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(CP_ERR_HIP, "batch failed");
    cp_batch_destroy(b);
    return rc;
}

// ------------------------------------------------------------------------------ autotune

__global__ void autotune_fill_kernel(uint8_t* data, uint32_t bps, uint64_t n_elems, uint64_t frame_elems,
                                     uint32_t maxv) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_elems;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t f = i / frame_elems, e = i - f * frame_elems;
        const uint64_t h = splitmix64(i * 0x9E3779B97F4A7C15ull);
        // per-element level + small noise, with one frame in eight replaced by a distant value
        int64_t v = (int64_t)(splitmix64(e) % (maxv / 2)) + (int64_t)(maxv / 4);
        v += (int64_t)(h & 15) - 8;
        if ((splitmix64(f * 131 + e / 97) & 7) == 0) v = (int64_t)(h >> 40) % maxv;
        if (v < 0) v = 0;
        if (v > (int64_t)maxv) v = maxv;
        if (bps == 1) data[i] = (uint8_t)v;
        else reinterpret_cast<uint16_t*>(data)[i] = (uint16_t)v;
    }
}

// Public encoding of a kernel choice: (kernel << 24) | (pixels per lane << 16) | (lanes per
// pixel << 8) | frames per lane; kernel 0: 16-bit halves, 1: 8-bit packed bytes with the byte
// search, 2: 8-bit packed bytes with sorting-network ranks.
static int32_t encode_choice(int choice) {
    if (choice >= kU8Base) {
        const U8Shape& u = kU8Shapes[choice - kU8Base];
        return ((u.S ? 2 : 1) << 24) | (u.X << 16) | (u.G << 8) | u.F;
    }
    return (kShapes[choice].X << 16) | (kShapes[choice].G << 8) | kShapes[choice].F;
}

extern "C" CP_API int cp_tile_shapes_set(cp_ctx* ctx, uint32_t bps, uint32_t channels, const int32_t* shapes) {
    if (!ctx || !shapes) return fail(CP_ERR_INVALID_ARG, "null argument");
    if ((bps != 1 && bps != 2) || channels < 1 || channels > 4)
        return fail(CP_ERR_INVALID_ARG, "tile shapes cover 8/16-bit stacks with 1 to 4 channels");
    int8_t pick[2][kNumBuckets];
    for (int kind = 0; kind < 2; ++kind)
        for (int b = 0; b < kNumBuckets; ++b) {
            const int32_t want = shapes[kind * kNumBuckets + b];
            pick[kind][b] = -1;
            if (want == 0) continue;
            for (int c = 0; c < kMaxCand; ++c) {
                const int ch = kBucketCand[bps - 1][channels >= 3][b][c];
                if (ch >= 0 && want == encode_choice(ch)) pick[kind][b] = (int8_t)ch;
            }
            if (pick[kind][b] < 0)
                return fail(CP_ERR_INVALID_ARG, "shape 0x%07x is not a candidate for frame bucket %d", want, b);
        }
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (int kind = 0; kind < 2; ++kind)
        for (int b = 0; b < kNumBuckets; ++b) ctx->tuned[bps - 1][channels - 1][kind][b] = pick[kind][b];
    return CP_OK;
}

extern "C" CP_API int cp_autotune(cp_ctx* ctx, uint32_t bps, uint32_t channels, int32_t* shapes) {
    if (!ctx) return fail(CP_ERR_INVALID_ARG, "null ctx");
    if ((bps != 1 && bps != 2) || channels < 1 || channels > 4)
        return fail(CP_ERR_INVALID_ARG, "autotune covers 8/16-bit stacks with 1 to 4 channels");
    // an image of 2048 x 96 pixels: every tile shape fills the device at least 12 times over
    const uint32_t W = 2048, H = 96, NMAX = kBucketMax[kNumBuckets - 1];
    cp_stack* st = nullptr;
    int rc = cp_stack_create(ctx, W, H, channels, bps, NMAX, &st);
    if (rc) return rc;
    const uint64_t fe = (uint64_t)W * H * channels;
    const uint64_t pos_bytes = 8ull * NMAX, out_bytes = fe * (bps + 1);
    uint64_t* d_pos = nullptr;
    void* d_out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool charged = false, ok = true;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        hipStream_t s = ctx->stream;
        rc = ctx->charge(pos_bytes + out_bytes);
        charged = rc == CP_OK;
        ok = charged;
        if (ok) {
            hipLaunchKernelGGL(autotune_fill_kernel, dim3(1024), dim3(256), 0, s, st->dev, bps, fe * NMAX, fe,
                               bps == 1 ? 255u : 65535u);
            st->count = NMAX;
        }
        std::vector<uint64_t> pos(NMAX);
        for (uint32_t t = 0; t < NMAX; ++t) pos[t] = (uint64_t)t * fe;
        ok = ok && hipMalloc((void**)&d_pos, pos_bytes) == hipSuccess;
        ok = ok && hipMalloc(&d_out, out_bytes) == hipSuccess;
        ok = ok && hipEventCreate(&e0) == hipSuccess;
        ok = ok && hipEventCreate(&e1) == hipSuccess;
        ok = ok && hipMemcpyAsync(d_pos, pos.data(), pos_bytes, hipMemcpyHostToDevice, s) == hipSuccess;
        auto run = [&](int choice, const OutlierK& k) {
            if (choice >= kU8Base) launch_outlier_u8(channels, k, choice - kU8Base, s);
            else launch_outlier_fast(bps, channels, k, choice, s);
        };
        for (int kb = 0; ok && kb < 2 * kNumBuckets; ++kb) {
            const int kind = kb / kNumBuckets, b = kb % kNumBuckets;  // kind 0: absolute, 1: relative
            const int* cand = kBucketCand[bps - 1][channels >= 3][b];
            int best = cand[0];
            if (cand[1] >= 0) {
                // two passes over the candidates, the second in reverse order, keeping each
                // candidate's faster time: a clock still ramping up in the first pass would
                // otherwise count against the candidates measured first
                int nc = 0;
                while (nc < kMaxCand && cand[nc] >= 0) ++nc;
                float cand_ms[kMaxCand];
                for (int c = 0; c < nc; ++c) cand_ms[c] = 1e30f;
                for (int step = 0; ok && step < 2 * nc; ++step) {
                    const int c = step < nc ? step : 2 * nc - 1 - step;
                    OutlierK k{};
                    k.stack = st->dev;
                    k.pos_off = d_pos;
                    k.n = kBucketMax[b];
                    k.n_sample = k.n;
                    k.stack_w = W;
                    k.out_w = W;
                    k.out_h = H;
                    const float range = bps == 1 ? 255.0f : 65535.0f;
                    k.th = kind == 0 ? ThresholdK{1, 0.05f * range, 0.2f * range, 1.0f / (0.15f * range)}
                                     : ThresholdK{0, 2.0f, 4.0f, 0.5f};
                    k.background = kind == 1 && bps == 1 ? 2 : 0;  // 16-bit average sums need N <= 256
                    k.outlier = kind == 1 && bps == 1 ? 3 : 2;
                    for (int i = 0; i < 4; ++i) k.w[i] = 1.0f;
                    k.fade.is_none = 1;
                    k.out_px = d_out;
                    k.out_mask = (uint8_t*)d_out + fe * bps;
                    run(cand[c], k);  // warm-up
                    ok = hipEventRecord(e0, s) == hipSuccess;
                    for (int r = 0; r < 3; ++r) run(cand[c], k);
                    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
                    float ms = 0.0f;
                    ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                    if (ok) cand_ms[c] = std::min(cand_ms[c], ms);
                }
                float best_ms = 1e30f;
                for (int c = 0; c < nc; ++c)
                    if (cand_ms[c] < best_ms) {
                        best_ms = cand_ms[c];
                        best = cand[c];
                    }
            }
            if (!ok) break;
            ctx->tuned[bps - 1][channels - 1][kind][b] = (int8_t)best;
            if (shapes) shapes[kb] = best >= 0 ? encode_choice(best) : 0;
        }
        (void)hipStreamSynchronize(s);
        if (d_pos) (void)hipFree(d_pos);
        if (d_out) (void)hipFree(d_out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (charged) ctx->uncharge(pos_bytes + out_bytes);
    }
    cp_stack_destroy(st);
    if (!charged) return rc;
    if (!ok) return fail(CP_ERR_HIP, "autotune failed: %s", hipGetErrorString(hipGetLastError()));
    return CP_OK;
}

