"""Serving loop: requests admitted while others are in flight give the same bytes as the same
composite run alone; admission respects the HBM budget."""
import numpy as np
import pytest

import chrono_photo_amd as cpa
import oracle_py as O

pytestmark = pytest.mark.gpu


def video_scene(n=40, w=96, h=64):
    return cpa.Scene(seed=11, width=w, height=h, n_frames=n, n_disks=2, disk_r_min=5, disk_r_max=9,
                     blob_permille=100, blob_r=6)


def test_windows_and_uploads_in_the_mix_equal_alone(ctx):
    sc = video_scene()
    frames = sc.generate()
    video = cpa.Stack.from_array(ctx, frames)
    wins = cpa.video_windows(cpa.FrameRange(0, 12, 1), cpa.FrameRange(0, 26, 1), sc.n_frames)
    proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), cpa.BackgroundMode.FIRST,
                                cpa.OutlierSelectionMode.FORWARD)
    rel = cpa.OutlierProcessor(cpa.Threshold.rel(2.0, 4.0), cpa.BackgroundMode.AVERAGE,
                               cpa.OutlierSelectionMode.AVERAGE)
    srv = cpa.Server(ctx, hbm_budget=256 << 20, max_inflight=6)
    tickets, expect = [], []
    rng = np.random.default_rng(5)
    for i, (f, idx) in enumerate(wins):
        job = cpa.Job(cpa.CP_MODE_OUTLIER, video, None, processor=proc, image_indices=idx)
        out = np.zeros((sc.height, sc.width, 3), np.uint8)
        mask = np.zeros_like(out)
        tickets.append(srv.submit(job, host_out=out, host_mask=mask))
        expect.append(("win", idx, out, mask))
        if i % 5 == 0:  # interleave single-image jobs that bring their own frames
            n = int(rng.integers(8, 30))
            x = cpa.Scene(seed=100 + i, width=80, height=48, n_frames=n, n_disks=1, disk_r_min=6,
                          disk_r_max=6).generate()
            p = rel if i % 10 == 0 else proc
            job = cpa.Job(cpa.CP_MODE_OUTLIER, None, None, processor=p, n_frames=n, bytes_per_sample=1)
            out = np.zeros((48, 80, 3), np.uint8)
            mask = np.zeros_like(out)
            tickets.append(srv.submit(job, host_frames=x, host_out=out, host_mask=mask))
            expect.append(("img", (x, p), out, mask))
    lat = [srv.wait(t) for t in tickets]
    assert all(v >= 0 for v in lat)
    st = srv.stats()
    assert st["completed"] == len(tickets) and st["peak_bytes"] <= 256 << 20
    # how many requests overlap depends on timing (small jobs may finish before the next
    # submit); the results above must not depend on it
    assert 1 <= st["peak_inflight"] <= 6
    srv.close()
    for kind, what, out, mask in expect:
        if kind == "win":
            ref_img, ref_mask = proc.process(video, image_indices=what)
        else:
            x, p = what
            st1 = cpa.Stack.from_array(ctx, x)
            ref_img, ref_mask = p.process(st1)
            st1.close()
        assert np.array_equal(out, ref_img), (kind,)
        assert np.array_equal(mask, ref_mask), (kind,)
    # one window against the oracle as well
    f, idx = wins[13]
    oimg, omask, _ = O.outlier(frames, O.threshold(True, 0.05, 0.2), 0, 4, slots=idx)
    assert np.array_equal(expect[[i for i, e in enumerate(expect) if e[0] == "win"][13]][2], oimg)


def test_pinned_results_while_uploads_are_in_flight(ctx):
    """Results written into pinned host buffers (by a copy kernel while uploads are in flight,
    else by DMA) equal the same composites run alone."""
    from chrono_photo_amd import hip as H
    sc = video_scene(n=30, w=320, h=240)
    video = cpa.Stack.from_array(ctx, sc.generate())
    wins = cpa.video_windows(cpa.FrameRange(0, 10, 1), cpa.FrameRange(0, 20, 1), sc.n_frames)
    proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), cpa.BackgroundMode.FIRST,
                                cpa.OutlierSelectionMode.EXTREME)
    big = cpa.Scene(seed=77, width=640, height=480, n_frames=64, n_disks=1, disk_r_min=9, disk_r_max=9).generate()
    frames = H.PinnedArray(big.shape, np.uint8)
    frames.array[...] = big
    srv = cpa.Server(ctx, max_inflight=4)
    tickets, outs = [], []
    for k in range(3):  # each upload of 59 MB stays in flight while windows start
        out = H.PinnedArray((480, 640, 3), np.uint8)
        job = cpa.Job(cpa.CP_MODE_OUTLIER, None, None, processor=proc, n_frames=64, bytes_per_sample=1)
        tickets.append(srv.submit(job, host_frames=frames.array, host_out=out.array))
        outs.append(("img", None, out))
        for f, idx in wins[7 * k:7 * k + 7]:
            o = H.PinnedArray((sc.height, sc.width, 3), np.uint8)
            m = H.PinnedArray((sc.height, sc.width, 3), np.uint8)
            job = cpa.Job(cpa.CP_MODE_OUTLIER, video, None, processor=proc, image_indices=idx)
            tickets.append(srv.submit(job, host_out=o.array, host_mask=m.array))
            outs.append(("win", idx, (o, m)))
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    srv.close()
    print("downloads by copy kernel:", st["kernel_downloads"])
    st_big = cpa.Stack.from_array(ctx, big)
    ref_big, _ = proc.process(st_big, with_mask=False)
    for kind, idx, bufs in outs:
        if kind == "img":
            assert np.array_equal(bufs.array, ref_big)
        else:
            ref_img, ref_mask = proc.process(video, image_indices=idx)
            assert np.array_equal(bufs[0].array, ref_img) and np.array_equal(bufs[1].array, ref_mask)
    st_big.close()
    video.close()


def test_budget_admission_and_rejection(ctx):
    sc = video_scene(n=20, w=64, h=48)
    x = sc.generate()
    per_request = x.nbytes + 64 * 48 * 3 + 64 * 48 * 3
    srv = cpa.Server(ctx, hbm_budget=per_request + 1000, max_inflight=4)  # one at a time fits
    proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), 0, 2)
    outs, tickets = [], []
    for _ in range(6):
        job = cpa.Job(cpa.CP_MODE_OUTLIER, None, None, processor=proc, n_frames=20, bytes_per_sample=1)
        out = np.zeros((48, 64, 3), np.uint8)
        mask = np.zeros_like(out)
        tickets.append(srv.submit(job, host_frames=x, host_out=out, host_mask=mask))
        outs.append((out, mask))
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    assert st["peak_bytes"] <= per_request + 1000 and st["peak_inflight"] == 1
    ref = O.outlier(x, O.threshold(True, 0.05, 0.2), 0, 2)
    for out, mask in outs:
        assert np.array_equal(out, ref[0]) and np.array_equal(mask, ref[1])
    big = np.zeros((200, 48, 64, 3), np.uint8)
    job = cpa.Job(cpa.CP_MODE_OUTLIER, None, None, processor=proc, n_frames=200, bytes_per_sample=1)
    t = srv.submit(job, host_frames=big, host_out=np.zeros((48, 64, 3), np.uint8))
    with pytest.raises(cpa.ChronoError) as e:
        srv.wait(t)
    assert e.value.code == cpa.CP_ERR_OUT_OF_HBM_BUDGET
    srv.close()
