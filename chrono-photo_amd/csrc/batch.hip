// batch.hip -- batched launches (cp_batch_*), the tile-shape autotuner and the serving loop.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "k_batch.h"
#include "kernels_u8.h"
#include "runtime.h"

using namespace cp;

namespace {

// A job that runs as its own launch keeps deep copies of the caller's arrays.
struct DirectJob {
    cp_job job;
    std::vector<uint32_t> frames;
    std::vector<int32_t> crop;
    std::vector<uint32_t> sample;
    std::vector<float> fade_outlier, fade_simple;
};

struct Group {
    uint32_t bps = 0, channels = 0;
    int shape = 0;
    bool absolute = false;
    std::vector<BatchJobK> jobs;      // host copies (device pointers patched at upload)
    std::vector<std::vector<uint64_t>> pos;
    std::vector<std::vector<float>> fades;
    std::vector<uint64_t> cost;
    uint32_t total = 0;
    BatchJobK* d_jobs = nullptr;
};

bool batch_fast_ok(const cp_stack* st, const OutlierPrep& p) {
    if (!p.sample_mask.empty()) return false;
    if (st->channels < 3 || st->bps != 1) return false;
    const bool sums = p.k.background == 2 || p.k.outlier == 3;
    if (sums && st->bps == 2 && p.k.n > 256) return false;
    return p.k.n <= 256;
}

int launch_group(const Group& g, hipStream_t s) {
    const uint32_t nj = (uint32_t)g.jobs.size();
    if (g.channels == 3) return launch_batch_b1_c3(g.d_jobs, nj, g.total, g.shape, g.absolute, s);
    return launch_batch_b1_c4(g.d_jobs, nj, g.total, g.shape, g.absolute, s);
}

}  // namespace

struct cp_batch {
    cp_ctx* ctx = nullptr;
    std::vector<Group> groups;
    std::vector<std::unique_ptr<DirectJob>> direct;
    void* d_tables = nullptr;
    uint64_t table_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int run_direct(cp_ctx* ctx, DirectJob& d) {
    cp_job& j = d.job;
    if (j.mode == CP_MODE_OUTLIER)
        return cp_outlier(ctx, j.stack, &j.frames, &j.outlier, &j.out, CP_FLAG_DEVICE_OUT);
    return cp_simple(ctx, j.stack, &j.frames, j.weights, &j.fade, j.mode == CP_MODE_DARKER, j.out.px,
                     CP_FLAG_DEVICE_OUT);
}

// Device resources of a batch (graph, staged tables); the context stream is drained first.
static void batch_release(cp_batch* b) {
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_tables) {
        (void)hipFree(b->d_tables);
        b->ctx->uncharge(b->table_bytes);
    }
    b->exec = nullptr;
    b->graph = nullptr;
    b->d_tables = nullptr;
    b->groups.clear();
    b->direct.clear();
}

void cp::batch_orphan(cp_batch* b) {
    batch_release(b);
    b->ctx = nullptr;
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
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const cp_job& j = jobs[i];
        if (!j.stack || !j.out.px) return fail(CP_ERR_INVALID_ARG, "job %u: null stack or output", i);
        if (j.stack->ctx != ctx) return fail(CP_ERR_INVALID_ARG, "job %u: stack belongs to another context", i);
        if (j.mode < CP_MODE_OUTLIER || j.mode > CP_MODE_DARKER) return fail(CP_ERR_INVALID_ARG, "job %u: bad mode", i);
        if (j.mode == CP_MODE_OUTLIER) {
            OutlierPrep prep;
            int rc = prepare_outlier(j.stack, &j.frames, &j.outlier, &prep);
            if (rc) return rc;
            int shape = ctx->choice_for(j.stack->bps, j.stack->channels, prep.k.n);
            if (shape < 0 || shape >= kU8Base) shape = fast_default_shape(j.stack->bps, prep.k.n);
            if (shape > kBatchShapeMax) shape = kBatchShapeMax;  // 8 x 32 covers up to 256 frames
            if (batch_fast_ok(j.stack, prep) && shape >= 0) {
                Group* g = nullptr;
                const bool absolute = prep.k.th.absolute != 0;
                for (auto& gg : b->groups)
                    if (gg.bps == j.stack->bps && gg.channels == j.stack->channels && gg.shape == shape &&
                        gg.absolute == absolute)
                        g = &gg;
                if (!g) {
                    b->groups.emplace_back();
                    g = &b->groups.back();
                    g->bps = j.stack->bps;
                    g->channels = j.stack->channels;
                    g->shape = shape;
                    g->absolute = absolute;
                }
                BatchJobK bj{};
                bj.k = prep.k;
                bj.k.out_px = j.out.px;
                bj.k.out_mask = j.out.mask;
                bj.k.warnings = (unsigned long long*)j.out.warnings;
                bj.k.stats = (PixelStats*)j.out.stats;
                bj.shape = shape;
                bj.tiles_x = (prep.k.out_w + (256 / kShapes[shape].G) - 1) / (256 / kShapes[shape].G);
                g->jobs.push_back(bj);
                g->pos.push_back(std::move(prep.pos_off));
                g->fades.push_back(std::move(prep.fade_values));
                g->cost.push_back((uint64_t)prep.k.n * prep.k.out_w * prep.k.out_h);
                continue;
            }
        }
        if (j.mode != CP_MODE_OUTLIER) {
            int rc = check_weights(j.weights);
            if (rc) return rc;
            Geometry geo;
            if ((rc = resolve_frames(j.stack, &j.frames, &geo))) return rc;
        }
        auto d = std::make_unique<DirectJob>();
        d->job = j;
        const uint32_t n = j.frames.n;
        if (j.frames.frames) {
            d->frames.assign(j.frames.frames, j.frames.frames + n);
            d->job.frames.frames = d->frames.data();
        }
        if (j.frames.crop_xy) {
            d->crop.assign(j.frames.crop_xy, j.frames.crop_xy + 2 * n);
            d->job.frames.crop_xy = d->crop.data();
        }
        if (j.outlier.sample_pos) {
            d->sample.assign(j.outlier.sample_pos, j.outlier.sample_pos + j.outlier.n_sample);
            d->job.outlier.sample_pos = d->sample.data();
        }
        if (!j.outlier.fade.is_none && j.outlier.fade.values) {
            d->fade_outlier.assign(j.outlier.fade.values, j.outlier.fade.values + j.outlier.fade.len);
            d->job.outlier.fade.values = d->fade_outlier.data();
        }
        if (!j.fade.is_none && j.fade.values) {
            d->fade_simple.assign(j.fade.values, j.fade.values + j.fade.len);
            d->job.fade.values = d->fade_simple.data();
        }
        b->direct.push_back(std::move(d));
    }

    // order each group's jobs by cost (largest first) and lay out its work list
    uint64_t bytes = 0;
    auto align = [](uint64_t v) { return (v + 255) & ~uint64_t(255); };
    for (auto& g : b->groups) {
        std::vector<size_t> ord(g.jobs.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) { return g.cost[a] > g.cost[c]; });
        std::vector<BatchJobK> jobs_sorted;
        std::vector<std::vector<uint64_t>> pos_sorted;
        std::vector<std::vector<float>> fade_sorted;
        for (size_t i : ord) {
            jobs_sorted.push_back(g.jobs[i]);
            pos_sorted.push_back(std::move(g.pos[i]));
            fade_sorted.push_back(std::move(g.fades[i]));
        }
        g.jobs.swap(jobs_sorted);
        g.pos.swap(pos_sorted);
        g.fades.swap(fade_sorted);
        uint64_t tiles = 0;
        for (auto& bj : g.jobs) {
            bj.tile0 = (uint32_t)tiles;
            bj.tiles = bj.tiles_x * bj.k.out_h;
            tiles += bj.tiles;
        }
        if (tiles >= 0x7FFFFFFFull) return fail(CP_ERR_CAPACITY, "batch has too many tiles");
        g.total = (uint32_t)tiles;
        bytes += align(sizeof(BatchJobK) * g.jobs.size());
        for (auto& p : g.pos) bytes += align(8 * p.size());
        for (auto& f : g.fades) bytes += align(4 * f.size());
    }
    if (bytes) {
        int rc = ctx->charge(bytes);
        if (rc) return rc;
        (void)hipSetDevice(ctx->device);
        if (hipMalloc(&b->d_tables, bytes) != hipSuccess) {
            ctx->uncharge(bytes);
            return fail(CP_ERR_HIP, "hipMalloc of %llu batch table bytes failed", (unsigned long long)bytes);
        }
        b->table_bytes = bytes;
        std::vector<uint8_t> host(bytes, 0);
        uint8_t* dbase = (uint8_t*)b->d_tables;
        uint64_t off = 0;
        for (auto& g : b->groups) {
            const uint64_t jobs_off = off;
            off += align(sizeof(BatchJobK) * g.jobs.size());
            for (size_t i = 0; i < g.jobs.size(); ++i) {
                memcpy(host.data() + off, g.pos[i].data(), 8 * g.pos[i].size());
                g.jobs[i].k.pos_off = (const uint64_t*)(dbase + off);
                off += align(8 * g.pos[i].size());
                if (!g.fades[i].empty()) {
                    memcpy(host.data() + off, g.fades[i].data(), 4 * g.fades[i].size());
                    g.jobs[i].k.fade.values = (const float*)(dbase + off);
                    off += align(4 * g.fades[i].size());
                }
            }
            memcpy(host.data() + jobs_off, g.jobs.data(), sizeof(BatchJobK) * g.jobs.size());
            g.d_jobs = (BatchJobK*)(dbase + jobs_off);
        }
        CP_HIP_TRY(hipMemcpy(b->d_tables, host.data(), bytes, hipMemcpyHostToDevice));
    }
    if (!(flags & CP_FLAG_NO_GRAPH) && !b->groups.empty()) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        CP_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int lrc = 0;
        for (auto& g : b->groups) lrc |= launch_group(g, ctx->stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
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
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        if (b->exec) {
            CP_HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
        } else {
            for (auto& g : b->groups)
                if (launch_group(g, ctx->stream))
                    return fail(CP_ERR_HIP, "batch kernel launch: %s", hipGetErrorString(hipGetLastError()));
        }
    }
    for (auto& d : b->direct) {
        int rc = run_direct(ctx, *d);
        if (rc) return rc;
    }
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

// Public encoding of a kernel choice: (pixels per lane << 16) | (lanes per pixel << 8) |
// frames per lane, with bit 24 set for the 8-bit packed-byte kernel.
static int32_t encode_choice(int choice) {
    if (choice >= kU8Base) {
        const U8Shape& u = kU8Shapes[choice - kU8Base];
        return (1 << 24) | (u.X << 16) | (u.G << 8) | u.F;
    }
    return (kShapes[choice].X << 16) | (kShapes[choice].G << 8) | kShapes[choice].F;
}

extern "C" CP_API int cp_tile_shapes_set(cp_ctx* ctx, uint32_t bps, uint32_t channels, const int32_t* shapes) {
    if (!ctx || !shapes) return fail(CP_ERR_INVALID_ARG, "null argument");
    if ((bps != 1 && bps != 2) || channels < 1 || channels > 4)
        return fail(CP_ERR_INVALID_ARG, "tile shapes cover 8/16-bit stacks with 1 to 4 channels");
    int8_t pick[kNumBuckets];
    for (int b = 0; b < kNumBuckets; ++b) {
        pick[b] = -1;
        if (shapes[b] == 0) continue;
        for (int c = 0; c < kMaxCand; ++c) {
            const int ch = kBucketCand[bps - 1][channels >= 3][b][c];
            if (ch >= 0 && shapes[b] == encode_choice(ch)) pick[b] = (int8_t)ch;
        }
        if (pick[b] < 0)
            return fail(CP_ERR_INVALID_ARG, "shape 0x%07x is not a candidate for frame bucket %d", shapes[b], b);
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (int b = 0; b < kNumBuckets; ++b) ctx->tuned[bps - 1][channels - 1][b] = pick[b];
    return CP_OK;
}

extern "C" CP_API int cp_autotune(cp_ctx* ctx, uint32_t bps, uint32_t channels, int32_t* shapes) {
    if (!ctx) return fail(CP_ERR_INVALID_ARG, "null ctx");
    if ((bps != 1 && bps != 2) || channels < 1 || channels > 4)
        return fail(CP_ERR_INVALID_ARG, "autotune covers 8/16-bit stacks with 1 to 4 channels");
    const uint32_t W = 2048, H = 24, NMAX = kBucketMax[kNumBuckets - 1];
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
        for (int b = 0; ok && b < kNumBuckets; ++b) {
            const int* cand = kBucketCand[bps - 1][channels >= 3][b];
            int best = cand[0];
            if (cand[1] >= 0) {
                float best_ms = 1e30f;
                for (int c = 0; ok && c < kMaxCand && cand[c] >= 0; ++c) {
                    OutlierK k{};
                    k.stack = st->dev;
                    k.pos_off = d_pos;
                    k.n = kBucketMax[b];
                    k.n_sample = k.n;
                    k.stack_w = W;
                    k.out_w = W;
                    k.out_h = H;
                    const float range = bps == 1 ? 255.0f : 65535.0f;
                    k.th = ThresholdK{1, 0.05f * range, 0.2f * range, 1.0f};
                    k.background = 0;
                    k.outlier = 2;
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
                    if (ok && ms < best_ms) {
                        best_ms = ms;
                        best = cand[c];
                    }
                }
            }
            if (!ok) break;
            ctx->tuned[bps - 1][channels - 1][b] = (int8_t)best;
            if (shapes) shapes[b] = best >= 0 ? encode_choice(best) : 0;
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

