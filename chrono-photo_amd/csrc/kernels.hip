// kernels.hip -- kernel instantiation and launch policy for the compositing modes.
#include "k_outlier_generic.h"
#include "k_simple.h"
#include "kernels_fast.h"
#include "kernels_u8.h"
#include "runtime.h"

namespace cp {

template <int B>
static int launch_outlier_generic_b(const cp_stack* st, const OutlierK& k, hipStream_t s) {
    dim3 block(256), grid((k.out_w + 255) / 256, k.out_h);
    switch (st->channels) {
    case 1: hipLaunchKernelGGL((outlier_generic_kernel<B, 1>), grid, block, 0, s, k); break;
    case 2: hipLaunchKernelGGL((outlier_generic_kernel<B, 2>), grid, block, 0, s, k); break;
    case 3: hipLaunchKernelGGL((outlier_generic_kernel<B, 3>), grid, block, 0, s, k); break;
    default: hipLaunchKernelGGL((outlier_generic_kernel<B, 4>), grid, block, 0, s, k); break;
    }
    return CP_OK;
}

OutlierPlan plan_outlier(cp_ctx* ctx, const cp_stack* st, const OutlierK& k, uint64_t max_pos, uint32_t flags) {
    // The register-resident kernels need exact integer sums for the average modes
    // (N * max sample < 2^24: always for 8-bit stacks).  8-bit composites take the packed-byte
    // kernel up to 1280 frames (and, with a `--sample` subset, up to 128 sample frames) or the
    // tuned choice between the two kernels; 16-bit ones the 16-bit kernel up to 512 frames
    // without a sample subset; the one-lane-per-pixel generic kernel takes the rest.  A
    // packed-byte choice whose wide loads would read past a borrowed stack's end falls back to
    // the bucket's 16-bit-halves shape, or to the generic kernel where the bucket has none.
    OutlierPlan p;
    const bool sums = k.background == 2 || k.outlier == 3;
    const bool exact_sums = st->bps == 1 || k.n <= 256;
    if ((flags & CP_FLAG_FORCE_GENERIC) || (sums && !exact_sums)) return p;
    auto u8_fits = [&](int sh) {
        return u8_tile_fits(st->dev, st->alloc_bytes, st->channels, kU8Shapes[sh].X, k, max_pos);
    };
    if (k.sample_mask != nullptr) {
        if (st->bps == 1 && k.n <= kU8MaxFrames && k.n_sample <= kU8MaxSample && k.samp_off) {
            const int sh = u8_shape_for(k.n, true, k.n_sample);
            if (sh >= 0 && u8_fits(sh)) p = {kPlanU8, sh};
        }
        return p;
    }
    int choice = ctx->choice_for(st->bps, st->channels, k.n, k.th.absolute != 0);
    if (choice >= kU8Base) {
        if (u8_fits(choice - kU8Base)) return {kPlanU8, choice - kU8Base};
        // a borrowed stack the wide loads would overrun: the bucket's 16-bit-halves shape
        const int b = bucket_for(k.n);
        choice = b >= 0 && kBucketCand[st->bps - 1][st->channels >= 3][b][0] < kU8Base
                     ? kBucketCand[st->bps - 1][st->channels >= 3][b][0]
                     : -1;
    }
    if (choice >= 0) {
        if (kShapes[choice].X == 2 && !pair_tile_fits(st->dev, st->alloc_bytes, st->bps, st->channels, k, max_pos))
            choice = narrow_shape(choice);
        if (choice >= 0) p = {kPlanFast, choice};
    }
    return p;
}

int launch_outlier(cp_ctx* ctx, const cp_stack* st, const OutlierK& k, uint64_t max_pos, uint32_t flags,
                   hipStream_t s) {
    const OutlierPlan p = plan_outlier(ctx, st, k, max_pos, flags);
    bool done = false;
    if (p.family == kPlanU8) done = launch_outlier_u8(st->channels, k, p.shape, s);
    else if (p.family == kPlanFast) done = launch_outlier_fast(st->bps, st->channels, k, p.shape, s);
    if (!done) {
        if (st->bps == 1) launch_outlier_generic_b<1>(st, k, s);
        else launch_outlier_generic_b<2>(st, k, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CP_ERR_HIP, "outlier kernel launch: %s", hipGetErrorString(e));
    return CP_OK;
}

template <int B, int PX>
static void launch_simple_b(const cp_stack* st, const SimpleK& k, hipStream_t s) {
    const uint32_t lanes = (k.out_w + PX - 1) / PX;
    dim3 block(256), grid((lanes + 255) / 256, k.out_h);
    switch (st->channels) {
    case 1: hipLaunchKernelGGL((simple_kernel<B, 1, PX>), grid, block, 0, s, k); break;
    case 2: hipLaunchKernelGGL((simple_kernel<B, 2, PX>), grid, block, 0, s, k); break;
    case 3: hipLaunchKernelGGL((simple_kernel<B, 3, PX>), grid, block, 0, s, k); break;
    default: hipLaunchKernelGGL((simple_kernel<B, 4, PX>), grid, block, 0, s, k); break;
    }
}

int launch_simple(cp_ctx* ctx, const cp_stack* st, const SimpleK& k, hipStream_t s) {
    (void)ctx;
    if (st->bps == 1) launch_simple_b<1, 4>(st, k, s);
    else launch_simple_b<2, 2>(st, k, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CP_ERR_HIP, "simple kernel launch: %s", hipGetErrorString(e));
    return CP_OK;
}

}  // namespace cp
