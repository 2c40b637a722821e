"""Batched launches (one launch per tile shape over the tiles of many stacks, captured in a HIP
graph) give byte-identical results to one launch per stack, and the autotuner's tile choice
never changes results."""
import numpy as np
import pytest

import chrono_photo_amd as cpa
from chrono_photo_amd import hip as H
import oracle_py as O

pytestmark = pytest.mark.gpu

SELS = list(cpa.OutlierSelectionMode)


def scene_stack(ctx, seed, n, w, h, c=3, bps=1):
    sc = cpa.Scene(seed=seed, width=w, height=h, channels=c, bytes_per_sample=bps, n_frames=n,
                   base_lo=40 if bps == 1 else 8000, base_hi=200 if bps == 1 else 50000,
                   sigma=3.0 if bps == 1 else 200.0, n_disks=2, disk_r_min=4, disk_r_max=9,
                   disk_val_lo=15 if bps == 1 else 2000, disk_val_hi=25 if bps == 1 else 3000,
                   blob_permille=100, blob_r=6, blob_val=230 if bps == 1 else 60000)
    x = sc.generate()
    return x, cpa.Stack.from_array(ctx, x)


def make_jobs(ctx, specs):
    jobs, bufs = [], []
    for (seed, n, w, h, c, bps, mode, absolute, sel, bg) in specs:
        x, st = scene_stack(ctx, seed, n, w, h, c, bps)
        px = H.DeviceBuffer(w * h * c * bps)
        mask = H.DeviceBuffer(w * h * c)
        warn = H.DeviceBuffer(8)
        warn.from_host(np.zeros(1, np.uint64))
        if mode == cpa.CP_MODE_OUTLIER:
            thr = cpa.Threshold.abs(0.05, 0.2) if absolute else cpa.Threshold.rel(2.0, 4.0)
            proc = cpa.OutlierProcessor(thr, bg, sel)
            job = cpa.Job(mode, st, px.ptr, mask.ptr, warn.ptr, processor=proc)
        else:
            job = cpa.Job(mode, st, px.ptr)
        jobs.append(job)
        bufs.append((x, st, px, mask, warn))
    return jobs, bufs


def read(bufs, job):
    x, st, px, mask, warn = bufs
    h, w, c = job.out_shape()
    img = px.to_host(np.zeros((h, w, c), x.dtype))
    m = mask.to_host(np.zeros((h, w, c), np.uint8))
    return img, m


def specs_mixed():
    rng = np.random.default_rng(3)
    specs = []
    for i in range(14):
        n = int(rng.integers(8, 200))
        w = int(rng.integers(40, 140))
        h = int(rng.integers(20, 60))
        specs.append((100 + i, n, w, h, 3 if i % 5 else 4, 1, cpa.CP_MODE_OUTLIER, i % 2 == 0,
                      SELS[i % 6], i % 4))
    specs.append((200, 40, 64, 32, 3, 2, cpa.CP_MODE_OUTLIER, True, SELS[4], 0))   # u16: own launch
    specs.append((201, 300, 48, 20, 3, 1, cpa.CP_MODE_OUTLIER, True, SELS[2], 0))  # > 256 frames
    specs.append((202, 30, 50, 30, 3, 1, cpa.CP_MODE_DARKER, True, SELS[0], 0))
    specs.append((203, 30, 50, 30, 4, 1, cpa.CP_MODE_LIGHTER, True, SELS[0], 0))
    return specs


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_plain_launches(ctx, graph):
    specs = specs_mixed()
    jobs, bufs = make_jobs(ctx, specs)
    batch = cpa.Batch(ctx, jobs, graph=graph)
    for rep in range(3):  # graph replays, with eager work between them
        for b in bufs:
            b[2].from_host(np.zeros(b[2].nbytes, np.uint8))
        batch.launch()
        ctx.synchronize()
        for spec, job, b in zip(specs, jobs, bufs):
            img, m = read(b, job)
            x, st = b[0], b[1]
            if spec[6] == cpa.CP_MODE_OUTLIER:
                ref_img, ref_mask = job.processor.process(st)
                assert np.array_equal(img, ref_img), (rep, spec)
                assert np.array_equal(m, ref_mask), (rep, spec)
            else:
                ref = cpa.SimpleProcessor(darker=spec[6] == cpa.CP_MODE_DARKER).process(st)
                assert np.array_equal(img, ref), (rep, spec)
    batch.close()


def test_batch_against_oracle(ctx):
    specs = specs_mixed()[:6]
    jobs, bufs = make_jobs(ctx, specs)
    cpa.batch_run(ctx, jobs)
    for spec, job, b in zip(specs, jobs, bufs):
        img, m = read(b, job)
        x = b[0]
        _, n, w, h, c, bps, mode, absolute, sel, bg = spec
        thr = O.threshold(absolute, 0.05 if absolute else 2.0, 0.2 if absolute else 4.0, bps)
        oimg, omask, _ = O.outlier(x, thr, bg, int(sel))
        assert np.array_equal(img, oimg), spec
        assert np.array_equal(m, omask), spec


# 8-bit 3-channel candidates per bucket: 16-bit-halves shapes (G, F, X), packed-byte byte-search
# (kernel 1) and sorting-network (kernel 2) shapes; each pins the group kernel the batch uses
CANDIDATES_U8 = [[(1, 16, 1), (1, 16, 1, 1), (1, 16, 1, 2)], [(1, 32, 1), (2, 16, 1), (1, 32, 1, 1)],
                 [(2, 32, 1), (4, 16, 1), (2, 32, 1, 1)], [(4, 32, 1), (8, 16, 1), (4, 32, 1, 1)],
                 [(8, 32, 1), (4, 64, 1), (8, 32, 1, 1)], [(8, 64, 1), (16, 32, 4, 1), (16, 32, 1, 1)]]


def test_batch_with_every_candidate_shape_against_oracle(ctx):
    """Batched groups under each pinned tile shape (cp_tile_shapes_set) against the oracle,
    on stacks whose first frames are a run of outliers longer than a 16-frame block."""
    rng = np.random.default_rng(9)
    try:
        for pick in (0, 1, 2):
            cpa.tile_shapes_set(ctx, 1, 3, [c[min(pick, len(c) - 1)] for c in CANDIDATES_U8])
            jobs, keep = [], []
            for i, n in enumerate((20, 40, 64, 100, 128, 200)):
                x = rng.integers(60, 70, (n, 6, 33, 3))
                x[: (n - 1) // 2, 2, 3:9, :] = 5
                hit = rng.random((n, 6, 33)) < 0.05
                x = np.where(hit[..., None], rng.integers(0, 256, x.shape), x).astype(np.uint8)
                st = cpa.Stack.from_array(ctx, x)
                px, mask = H.DeviceBuffer(x[0].size), H.DeviceBuffer(x[0].size)
                absolute = i % 2 == 0
                thr = cpa.Threshold.abs(0.05, 0.2) if absolute else cpa.Threshold.rel(2.0, 4.0)
                sel = SELS[(i + pick) % 6]
                proc = cpa.OutlierProcessor(thr, cpa.BackgroundMode.FIRST, sel)
                jobs.append(cpa.Job(cpa.CP_MODE_OUTLIER, st, px.ptr, mask.ptr, processor=proc))
                keep.append((x, st, px, mask, absolute, sel))
            cpa.batch_run(ctx, jobs)
            for job, (x, st, px, mask, absolute, sel) in zip(jobs, keep):
                img = px.to_host(np.zeros(x[0].shape, np.uint8))
                m = mask.to_host(np.zeros(x[0].shape, np.uint8))
                thr = O.threshold(absolute, 0.05 if absolute else 2.0, 0.2 if absolute else 4.0, 1)
                oimg, omask, _ = O.outlier(x, thr, 0, int(sel))
                assert np.array_equal(img, oimg), (pick, x.shape[0], absolute, sel)
                assert np.array_equal(m, omask), (pick, x.shape[0], absolute, sel)
                st.close()
                px.free()
                mask.free()
    finally:
        cpa.tile_shapes_set(ctx, 1, 3, [])


def test_batch_random_mixes(ctx):
    """Five batches of 12 jobs drawn from a fixed seed (outlier and darker/lighter, u8/u16,
    1-4 channels, 1-300 frames, 1x1 to 150x40 images, every mode), graph-replayed twice, each
    job against the oracle."""
    rng = np.random.default_rng(20261207)
    for batch_no in range(5):
        jobs, keep = [], []
        for j in range(12):
            bps = int(rng.choice([1, 2]))
            c = int(rng.integers(1, 5))
            n = int(rng.choice([int(rng.integers(1, 40)), int(rng.integers(40, 300))]))
            w, h = int(rng.integers(1, 150)), int(rng.integers(1, 40))
            hi = (1 << (8 * bps)) - 1
            x = rng.integers(hi // 4, 3 * hi // 4, (1, h, w, c)) + rng.integers(-hi // 40, hi // 40 + 1, (n, h, w, c))
            hit = rng.random((n, h, w)) < 0.15
            x = np.where(hit[..., None], rng.integers(0, hi + 1, (n, h, w, c)), x)
            x = np.clip(x, 0, hi).astype(np.uint8 if bps == 1 else np.uint16)
            st = cpa.Stack.from_array(ctx, x)
            px, mask = H.DeviceBuffer(x[0].nbytes), H.DeviceBuffer(x[0].size)
            mode = int(rng.choice([cpa.CP_MODE_OUTLIER, cpa.CP_MODE_OUTLIER, cpa.CP_MODE_DARKER, cpa.CP_MODE_LIGHTER]))
            absolute = n < 3 or bool(rng.random() < 0.5)
            bg, sel = int(rng.integers(0, 4)), int(rng.integers(0, 6))
            if mode == cpa.CP_MODE_OUTLIER:
                thr = cpa.Threshold.abs(0.05, 0.2) if absolute else cpa.Threshold.rel(2.0, 4.0)
                proc = cpa.OutlierProcessor(thr, bg, sel, seed=7)
                jobs.append(cpa.Job(mode, st, px.ptr, mask.ptr, processor=proc))
            else:
                jobs.append(cpa.Job(mode, st, px.ptr))
            keep.append((x, st, px, mask, mode, absolute, bg, sel))
        batch = cpa.Batch(ctx, jobs, graph=True)
        for rep in range(2):
            for k in keep:
                k[2].from_host(np.zeros(k[2].nbytes, np.uint8))
            batch.launch()
            ctx.synchronize()
            for j, (x, st, px, mask, mode, absolute, bg, sel) in enumerate(keep):
                img = px.to_host(np.zeros(x[0].shape, x.dtype))
                if mode == cpa.CP_MODE_OUTLIER:
                    thr = O.threshold(absolute, 0.05 if absolute else 2.0, 0.2 if absolute else 4.0, x.dtype.itemsize)
                    oimg, omask, _ = O.outlier(x, thr, bg, sel, seed=7)
                    m = mask.to_host(np.zeros(x[0].shape, np.uint8))
                    assert np.array_equal(img, oimg) and np.array_equal(m, omask), (batch_no, j, rep, x.shape, bg, sel)
                else:
                    ref = O.simple(x, darker=mode == cpa.CP_MODE_DARKER)
                    assert np.array_equal(img, ref), (batch_no, j, rep, x.shape, mode)
        batch.close()
        for x, st, px, mask, *_ in keep:
            st.close()
            px.free()
            mask.free()


@pytest.mark.parametrize("bps", [1, 2])
def test_autotune_keeps_results(ctx, bps):
    tuned = cpa.autotune(ctx, bps, 3)
    assert set(tuned) == {"abs", "rel"}
    for sh, nmax in zip(tuned["abs"] + tuned["rel"], cpa.TILE_BUCKETS * 2):
        if bps == 2 and nmax > 512:
            assert sh is None
            continue
        g, f, xp, kind = sh
        assert g * f >= nmax and g in (1, 2, 4, 8, 16, 32) and xp in ((1, 4) if bps == 1 else (1, 2))
        assert kind in ((0, 1, 2) if bps == 1 else (0,))
    for n in (12, 30, 60, 120, 250) + ((700, 1200) if bps == 1 else ()):
        x, st = scene_stack(ctx, 7 + n, n, 70, 24, 3, bps)
        for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3)):
            proc = cpa.OutlierProcessor(cpa.Threshold(absolute, lo, hi), bg, sel)
            img, mask = proc.process(st)
            oimg, omask, _ = O.outlier(x, O.threshold(absolute, lo, hi, bps), bg, sel)
            assert np.array_equal(img, oimg), (n, absolute)
            assert np.array_equal(mask, omask), (n, absolute)


def test_batch_groups_and_side_branches(ctx):
    """One batch mixing every path: group launches for 8-bit 1/3/4-channel stacks (one with
    thousands of tiles), a sorting-network choice, `--sample` jobs, tall stacks and 16-bit
    two-pixel tiles (one with an odd row width, so the staged offsets are realigned row by
    row); darker and a borrowed stack that ends at its last frame beside them; three graph
    replays, every job against the oracle."""
    rng = np.random.default_rng(77)
    keep, jobs = [], []
    try:
        cpa.tile_shapes_set(ctx, 1, 3, [None, (1, 32, 1, 2)], [None, (1, 32, 1, 2)])
        specs = [(1, 3, 20, 300, 40), (1, 3, 30, 64, 33), (1, 4, 200, 120, 9), (1, 3, 500, 40, 6), (1, 3, 700, 31, 5),
                 (2, 3, 40, 50, 10), (1, 1, 90, 45, 7), (1, 3, 8, 2000, 400), (2, 3, 100, 51, 12),
                 (2, 4, 30, 33, 7)]
        for i, (bps, c, n, w, h) in enumerate(specs):
            hi = (1 << (8 * bps)) - 1
            x = rng.integers(hi // 4, 3 * hi // 4, (1, h, w, c)) + rng.integers(-hi // 40, hi // 40 + 1, (n, h, w, c))
            hit = rng.random((n, h, w)) < 0.1
            x = np.clip(np.where(hit[..., None], rng.integers(0, hi + 1, (n, h, w, c)), x), 0, hi)
            x = x.astype(np.uint8 if bps == 1 else np.uint16)
            st = cpa.Stack.from_array(ctx, x)
            absolute = i % 2 == 0
            thr = cpa.Threshold.abs(0.05, 0.2) if absolute else cpa.Threshold.rel(2.0, 4.0)
            sample = sorted(rng.choice(n, min(n, 50), replace=False).tolist()) if i in (1, 4) else None
            proc = cpa.OutlierProcessor(thr, i % 4, (i * 5) % 6, sample_indices=sample, seed=3)
            px, mask = H.DeviceBuffer(x[0].nbytes), H.DeviceBuffer(x[0].size)
            jobs.append(cpa.Job(cpa.CP_MODE_OUTLIER, st, px.ptr, mask.ptr, processor=proc))
            keep.append(("outlier", x, st, px, mask, absolute, i % 4, (i * 5) % 6, sample))
        x = rng.integers(0, 256, (25, 9, 40, 3)).astype(np.uint8)
        st = cpa.Stack.from_array(ctx, x)
        px = H.DeviceBuffer(x[0].nbytes)
        jobs.append(cpa.Job(cpa.CP_MODE_DARKER, st, px.ptr))
        keep.append(("darker", x, st, px, None, None, None, None, None))
        # a borrowed stack that ends at its last frame: the wide-load tiles would read past it
        xw = np.ascontiguousarray(rng.integers(0, 256, (40, 5, 27, 3)).astype(np.uint8))
        buf = H.DeviceBuffer(xw.nbytes)
        buf.from_host(xw)
        stw = cpa.Stack(ctx, 27, 5, 3, 1, device_ptr=buf.ptr, n_frames=40)
        px = H.DeviceBuffer(xw[0].nbytes)
        proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), 0, 2)
        jobs.append(cpa.Job(cpa.CP_MODE_OUTLIER, stw, px.ptr, None, processor=proc))
        keep.append(("outlier", xw, stw, px, None, True, 0, 2, None))
        batch = cpa.Batch(ctx, jobs, graph=True)
        info = batch.info()
        # every spec job has a register-resident plan; darker is launched on its own
        assert info["group_jobs"] >= len(specs) and info["direct_jobs"] >= 1, info
        assert info["group_jobs"] + info["direct_jobs"] == len(jobs), info
        for rep in range(3):
            for k in keep:
                k[3].from_host(np.zeros(k[3].nbytes, np.uint8))
            batch.launch()
            ctx.synchronize()
            for j, (kind, x, st, px, mask, absolute, bg, sel, sample) in enumerate(keep):
                img = px.to_host(np.zeros(x[0].shape, x.dtype))
                if kind == "darker":
                    assert np.array_equal(img, O.simple(x, darker=True)), (rep, j)
                    continue
                thr = O.threshold(absolute, 0.05 if absolute else 2.0, 0.2 if absolute else 4.0, x.dtype.itemsize)
                oimg, omask, _ = O.outlier(x, thr, bg, sel, seed=3, sample_pos=sample)
                assert np.array_equal(img, oimg), (rep, j, x.shape)
                if mask is not None:
                    assert np.array_equal(mask.to_host(np.zeros(x[0].shape, np.uint8)), omask), (rep, j)
        batch.close()
        buf.free()
    finally:
        cpa.tile_shapes_set(ctx, 1, 3, [])
        for k in keep:
            k[2].close()
            k[3].free()
            if k[4] is not None:
                k[4].free()
