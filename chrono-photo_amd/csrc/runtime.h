// runtime.h -- host-side context, stack and parameter staging shared by the ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/chrono_hip.h"
#include "kparams.h"

namespace cp {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define CP_HIP_TRY(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::cp::fail(CP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Ring of small per-call parameter blocks: pinned host staging + device mirror, copied
// asynchronously on the owning stream.  When the ring wraps the stream is drained first,
// so a block is never rewritten while a queued copy or kernel may still read it.
struct ParamArena {
    uint8_t* host = nullptr;
    uint8_t* dev = nullptr;
    size_t cap = 0, head = 0;
    int init(size_t bytes);
    void release();
    // returns device pointer of a copy of `data`, or nullptr on failure
    void* put(const void* data, size_t bytes, hipStream_t s);
};

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace cp

struct cp_ctx {
    // Handles created on this context.  cp_ctx_destroy releases their device resources first
    // and leaves each as an inert shell (ctx == nullptr) whose own destroy call only frees it,
    // so a caller may destroy the context before its stacks, batches or servers.
    std::mutex reg_mu;
    std::unordered_set<cp_stack*> stacks;
    std::unordered_set<cp_batch*> batches;
    std::unordered_set<cp_server*> servers;
    void attach(cp_stack* s) { std::lock_guard<std::mutex> l(reg_mu); stacks.insert(s); }
    void attach(cp_batch* b) { std::lock_guard<std::mutex> l(reg_mu); batches.insert(b); }
    void attach(cp_server* v) { std::lock_guard<std::mutex> l(reg_mu); servers.insert(v); }
    void detach(cp_stack* s) { std::lock_guard<std::mutex> l(reg_mu); stacks.erase(s); }
    void detach(cp_batch* b) { std::lock_guard<std::mutex> l(reg_mu); batches.erase(b); }
    void detach(cp_server* v) { std::lock_guard<std::mutex> l(reg_mu); servers.erase(v); }

    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;
    int cu_count = 0;
    cp::ParamArena arena;
    cp::DeviceBuffer out_px, out_mask, out_stats, warn;  // host-output staging
    std::mutex mu;
    // autotuned kernel choice per (sample bytes - 1, channels - 1, threshold kind (0: absolute,
    // 1: relative), frame bucket); -1 = default
    int8_t tuned[2][4][2][16];
    // the register-resident kernel choice for n frames (kernels_fast.h), or -1: none
    int choice_for(uint32_t bps, uint32_t channels, uint32_t n, bool absolute) const;
    int ensure(cp::DeviceBuffer& b, size_t bytes);
    int charge(uint64_t bytes);
    void uncharge(uint64_t bytes) { used = used >= bytes ? used - bytes : 0; }
};

// Bytes allocated past the last frame of an owned stack: the two-pixel outlier tiles read up
// to 16 bytes from a pixel pair's first sample.
constexpr uint64_t kStackSlack = 256;

struct cp_stack {
    cp_ctx* ctx = nullptr;  // nullptr once the context was destroyed (an inert shell)
    uint32_t width = 0, height = 0, channels = 0, bps = 0;
    uint32_t capacity = 0, count = 0;
    uint8_t* dev = nullptr;
    uint64_t alloc_bytes = 0;  // readable bytes from dev (owned stacks carry kStackSlack more)
    bool owned = false;
    uint64_t frame_bytes() const { return (uint64_t)width * height * channels * bps; }
    uint64_t frame_elems() const { return (uint64_t)width * height * channels; }
};

namespace cp {

// Resolved geometry of one composite (shared by every mode).
struct Geometry {
    uint32_t n = 0;
    uint32_t out_w = 0, out_h = 0;
    std::vector<uint64_t> pos_off;  // element offset per frame position
};

int resolve_frames(const cp_stack* st, const cp_frames* fr, Geometry* g);
int check_weights(const float w[4]);

// Validated arguments of one outlier composite; device pointers (pos_off, sample_mask,
// fade.values, outputs) are left for the caller to place.
struct OutlierPrep {
    OutlierK k;
    std::vector<uint64_t> pos_off;
    std::vector<uint32_t> sample_mask;  // empty: every frame is in the median sample
    std::vector<uint32_t> sample_pos;   // the sample positions, sorted (empty: all)
    std::vector<float> fade_values;     // empty: no fade
};
int prepare_outlier(const cp_stack* st, const cp_frames* frames, const cp_outlier_params* p, OutlierPrep* out);
// Stages an outlier composite's per-call tables in `arena` on stream s -- the frame offsets
// `pos` (the composite's, or a band stack's), the sample mask, the sample positions' offsets
// taken from `pos`, the fade table -- and points k at the device copies.
int stage_outlier(ParamArena& arena, const OutlierPrep& prep, const std::vector<uint64_t>& pos, hipStream_t s,
                  OutlierK* k);
FadeK fade_to_device(cp_ctx* ctx, const cp_fade* f, int* rc);

// The kernel an outlier composite runs on: the one-lane-per-pixel generic kernel, a shape of
// the register-resident 16-bit-halves kernel (kShapes index) or of the packed-byte kernel
// (kU8Shapes index).
enum PlanFamily { kPlanGeneric = 0, kPlanFast = 1, kPlanU8 = 2 };
struct OutlierPlan {
    int family = kPlanGeneric;
    int shape = -1;
};
// max_pos: the largest element offset in the composite's frame table.
OutlierPlan plan_outlier(cp_ctx* ctx, const cp_stack* st, const OutlierK& k, uint64_t max_pos, uint32_t flags);

// Kernel launchers (defined in the kernel translation units).
int launch_outlier(cp_ctx* ctx, const cp_stack* st, const OutlierK& k, uint64_t max_pos, uint32_t flags,
                   hipStream_t s);
int launch_simple(cp_ctx* ctx, const cp_stack* st, const SimpleK& k, hipStream_t s);

// Release a handle's device resources when its context is destroyed first (batch.hip,
// server.hip); the handle stays allocated, with ctx == nullptr, until its own destroy call.
void batch_orphan(cp_batch* b);
void server_orphan(cp_server* v);

// Looks up, once per device, the device address of cp_shake_noalloc's static scratch array
// (shake.hip); called by cp_ctx_create with `device` current.
int shake_scratch_init(int device);

}  // namespace cp
