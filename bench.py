"""Throughput of the per-pixel stack-compositing hot path on one MI355X.

Default workload (outlier mode on a 256-frame 4K RGB16 stack, shake-corrected):
  * 256 synthetic RGB16 frames 3840x2160 (gradient 8000->50000, noise sigma 200, three
    moving dark disks, a static world-space block texture, a seeded integer translation of
    up to +-8 px per frame), generated directly in HBM;
  * shake analysis with 5 anchors, anchor radius 32, search radius 12, then Crop::create;
  * the timed step: one outlier composite (threshold abs/0.05/0.2, background first,
    --outlier forward) over the cropped stack, image + blend mask written to HBM.

One step = one pass of the hot path over the resident stack.  pixel-samples/s =
frames * out_h * out_w / step time.  The JSON line carries the roofline of the outlier
kernel (algorithmic bytes = stack read once + image + mask, over its average launch time
measured with HIP events on the launch stream) and the CPU oracle timed on the host on a
bounded band of output rows.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "chrono-photo_amd")]

HBM_PEAK_GBS = 8000.0


def measured_traffic(workload):
    """HBM bytes per launch of the workload's dominant kernel, as recorded in its committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE passes scaled by a plain copy's counts):
    the `traffic_bytes_per_launch:` line of profiles/<kernel>__<workload>.md, else None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(f"__{workload}.md") or name.startswith("shake_"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return {"bytes": int(line.split(":")[1]), "source": f"profiles/{name}"}
    return None


WORKLOADS = {
    "outlier_4k_rgb16_256f": dict(
        scene=dict(seed=2, width=3840, height=2160, channels=3, bytes_per_sample=2, n_frames=256,
                   base_lo=8000, base_hi=50000, sigma=200.0, n_disks=3, disk_r_min=20, disk_r_max=40,
                   disk_val_lo=2000, disk_val_hi=4000, shake_max=8, texture_amp=3000,
                   disk_y_lo=540, disk_y_hi=1620),
        # anchors on static texture outside the band the disks move in
        shake=dict(anchors=[(480, 150), (1920, 150), (3360, 150), (960, 2010), (2880, 2010)], r=32, s=12),
        mode="outlier", threshold=("abs", 0.05, 0.2), background="first", outlier="forward"),
    "outlier_1080p_rgb8_64f": dict(
        scene=dict(seed=1, width=1920, height=1080, channels=3, bytes_per_sample=1, n_frames=64,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=3, disk_r_min=10, disk_r_max=40,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=100, blob_r=30, blob_val=230),
        shake=None, mode="outlier", threshold=("rel", 2.0, 4.0), background="first", outlier="average"),
    "darker_1080p_rgb8_64f": dict(
        scene=dict(seed=1, width=1920, height=1080, channels=3, bytes_per_sample=1, n_frames=64,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=3, disk_r_min=10, disk_r_max=40,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=100, blob_r=30, blob_val=230),
        shake=None, mode="darker"),
    "lighter_1080p_rgb8_64f": dict(
        scene=dict(seed=1, width=1920, height=1080, channels=3, bytes_per_sample=1, n_frames=64,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=3, disk_r_min=10, disk_r_max=40,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=100, blob_r=30, blob_val=230),
        shake=None, mode="lighter"),
    "darker_4k_rgb16_256f": dict(
        scene=dict(seed=2, width=3840, height=2160, channels=3, bytes_per_sample=2, n_frames=256,
                   base_lo=8000, base_hi=50000, sigma=200.0, n_disks=3, disk_r_min=20, disk_r_max=40,
                   disk_val_lo=2000, disk_val_hi=4000, shake_max=8, texture_amp=3000,
                   disk_y_lo=540, disk_y_hi=1620),
        shake=None, mode="darker"),
    # a gallery-sized photo: about 1200 FHD frames per composite
    "outlier_1080p_rgb8_1200f": dict(
        scene=dict(seed=6, width=1920, height=1080, channels=3, bytes_per_sample=1, n_frames=1200,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=3, disk_r_min=10, disk_r_max=40,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=20, blob_r=30, blob_val=230),
        shake=None, mode="outlier", threshold=("abs", 0.05, 0.2), background="first", outlier="extreme"),
    # the same stack with `--sample 64`: medians over 64 seeded frame positions, distances over all
    "outlier_1080p_rgb8_1200f_s64": dict(
        scene=dict(seed=6, width=1920, height=1080, channels=3, bytes_per_sample=1, n_frames=1200,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=3, disk_r_min=10, disk_r_max=40,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=20, blob_r=30, blob_val=230),
        shake=None, mode="outlier", threshold=("abs", 0.05, 0.2), background="first", outlier="extreme",
        sample=64),
    "outlier_640x480_rgb8_16f": dict(
        scene=dict(seed=0, width=640, height=480, channels=3, bytes_per_sample=1, n_frames=16,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=1, disk_r_min=20, disk_r_max=20,
                   disk_val_lo=15, disk_val_hi=25),
        shake=None, mode="outlier", threshold=("abs", 0.05, 0.2), background="first", outlier="extreme"),
}


def _stack_from_scene(cpa, ctx, sc):
    st = cpa.Stack(ctx, sc.width, sc.height, sc.channels, sc.bytes_per_sample, capacity=sc.n_frames)
    sc.generate_device(st.device_ptr, 0, sc.n_frames, stream=ctx.stream)
    ctx.synchronize()
    st.commit(sc.n_frames)
    return st


BATCH_SETS = {
    # stacks that each fill the GPU on their own: batching can only save the gaps between launches
    "large": dict(n_stacks=48, frames=(8, 200), width=(640, 1920), height=(480, 1080), seed=3),
    # previews / thumbnails: each plain launch is bound by launch and synchronisation latency
    "small": dict(n_stacks=256, frames=(8, 64), width=(160, 640), height=(120, 480), seed=5),
}


def run_batching(cpa, H, ctx, kind="large"):
    """Independent RGB8 stacks of uneven sizes (abs/rel thresholds alternating, the six
    --outlier modes round-robin): one graph replay of the batch (one launch per kernel shape the
    jobs are planned on, covering all their tiles) vs one launch per stack, each waited for; the
    same launches queued back to back without waits are timed too, to show what the waits cost.
    "large": 48 stacks, frames U{8..200}, 640x480..1920x1080; "small": 256 stacks, frames
    U{8..64}, 160x120..640x480."""
    import numpy as np
    cfg = BATCH_SETS[kind]
    n_stacks = cfg["n_stacks"]
    rng = np.random.default_rng(cfg["seed"])
    jobs, bufs, total = [], [], 0
    sels = list(cpa.OutlierSelectionMode)
    for i in range(n_stacks):
        n = int(rng.integers(cfg["frames"][0], cfg["frames"][1] + 1))
        w = int(rng.integers(cfg["width"][0], cfg["width"][1] + 1))
        h = int(rng.integers(cfg["height"][0], cfg["height"][1] + 1))
        sc = cpa.Scene(seed=1000 + i, width=w, height=h, n_frames=n, n_disks=2, disk_r_min=10,
                       disk_r_max=30, blob_permille=100)
        st = _stack_from_scene(cpa, ctx, sc)
        thr = cpa.Threshold.abs(0.05, 0.2) if i % 2 == 0 else cpa.Threshold.rel(2.0, 4.0)
        proc = cpa.OutlierProcessor(thr, cpa.BackgroundMode.FIRST, sels[i % 6])
        px, mask = H.DeviceBuffer(w * h * 3), H.DeviceBuffer(w * h * 3)
        jobs.append(cpa.Job(cpa.CP_MODE_OUTLIER, st, px.ptr, mask.ptr, processor=proc))
        bufs.append((st, px, mask))
        total += n * w * h
    lib = ctx.lib

    def plain():
        for j in jobs:
            outs = cpa.cp_outputs(j.c.out.px, j.c.out.mask, None, None)
            cpa._check(lib.cp_outlier(ctx.handle, j.stack.handle, C.byref(j.sel.c), C.byref(j.c.outlier),
                                      C.byref(outs), cpa.CP_FLAG_DEVICE_OUT))
            ctx.synchronize()  # one stack per launch, each waited for like a drop-in call

    def timed(fn, reps=3):
        fn()
        ctx.synchronize()
        best = 1e30
        for _ in range(reps):
            e0, e1 = H.Event(), H.Event()
            t0 = time.perf_counter()
            e0.record(ctx.stream)
            fn()
            e1.record(ctx.stream)
            e1.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    def queued():
        for j in jobs:
            outs = cpa.cp_outputs(j.c.out.px, j.c.out.mask, None, None)
            cpa._check(lib.cp_outlier(ctx.handle, j.stack.handle, C.byref(j.sel.c), C.byref(j.c.outlier),
                                      C.byref(outs), cpa.CP_FLAG_DEVICE_OUT))

    plain_s = timed(plain)
    queued_s = timed(queued)
    ref = [b[1].to_host(np.zeros(b[1].nbytes, np.uint8)) for b in bufs]
    for b in bufs:  # the batched run must write every byte again
        b[1].from_host(np.zeros(b[1].nbytes, np.uint8))
    batch = cpa.Batch(ctx, jobs, graph=True)
    info = batch.info()
    batch_s = timed(lambda: batch.launch())
    same = all(np.array_equal(b[1].to_host(np.zeros(b[1].nbytes, np.uint8)), r) for b, r in zip(bufs, ref))
    batch.close()
    for st, px, mask in bufs:
        st.close()
        px.free()
        mask.free()
    return {"stacks": n_stacks, "frames": list(cfg["frames"]),
            "sizes": f"{cfg['width'][0]}x{cfg['height'][0]}..{cfg['width'][1]}x{cfg['height'][1]}",
            "pixel_samples": total, "plain_ms": plain_s * 1e3, "batched_ms": batch_s * 1e3,
            "queued_ms": queued_s * 1e3,
            "plain_value": total / plain_s, "batched_value": total / batch_s,
            "speedup": plain_s / batch_s, "speedup_vs_queued": queued_s / batch_s,
            "byte_identical": bool(same), "launches": info,
            "plain": f"{n_stacks} sequential cp_outlier launches, one stack each, each waited for",
            "queued": f"the same {n_stacks} cp_outlier launches queued back to back, one wait at the end",
            "batched": "one graph replay: a launch per kernel shape over all its jobs' tiles, "
                       "spread over four streams (cp_batch_create/cp_batch_launch), kernel choices "
                       "from cp_autotune"}


def run_workload(cpa, H, ctx, name, steps=10, warmup=2):
    """One outlier workload without shake correction (its stack generated in HBM), timed like
    the headline: HIP events around `steps` cp_outlier launches; its roofline from the same
    algorithmic bytes and its committed traffic summary."""
    import numpy as np
    wl = WORKLOADS[name]
    scene = cpa.Scene(**wl["scene"])
    n, w, h, ch, bps = scene.n_frames, scene.width, scene.height, scene.channels, scene.bytes_per_sample
    stack = _stack_from_scene(cpa, ctx, scene)
    sel = cpa._FrameSel(stack, None, None, n)
    kind, lo, hi = wl["threshold"]
    sample_pos = None
    if wl.get("sample"):
        sample_pos = np.sort(np.random.default_rng(7).choice(n, wl["sample"], replace=False))
    proc = cpa.OutlierProcessor(cpa.Threshold(kind == "abs", lo, hi), cpa.BackgroundMode.from_str(wl["background"]),
                                cpa.OutlierSelectionMode.from_str(wl["outlier"]), sample_indices=sample_pos)
    params = proc.params(bps)
    out_px, out_mask = H.DeviceBuffer(w * h * ch * bps), H.DeviceBuffer(w * h * ch)
    outs = cpa.cp_outputs(out_px.ptr, out_mask.ptr, None, None)
    lib = ctx.lib

    def step():
        cpa._check(lib.cp_outlier(ctx.handle, stack.handle, C.byref(sel.c), C.byref(params), C.byref(outs),
                                  cpa.CP_FLAG_DEVICE_OUT))
    for _ in range(warmup):
        step()
    ctx.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record(ctx.stream)
    for _ in range(steps):
        step()
    e1.record(ctx.stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    alg = n * w * h * ch * bps + w * h * ch * (bps + 1)
    gbs = alg / (ms / 1e3) / 1e9
    traffic = measured_traffic(name)
    stack.close()
    out_px.free()
    out_mask.free()
    return {"ms_per_step": ms, "value": n * w * h / (ms / 1e3), "unit": "pixel-samples/s",
            "config": {"frames": n, "width": w, "height": h, "channels": ch, "bytes_per_sample": bps,
                       "threshold": "/".join(map(str, wl["threshold"])), "background": wl["background"],
                       "outlier": wl["outlier"], **({"sample": wl["sample"]} if wl.get("sample") else {})},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                         **({"traffic_source": traffic["source"]} if traffic else {}),
                         "algorithmic_bytes_per_launch": alg}}


def run_serving(cpa, H, ctx, seed=4, n_jobs=24, window=60, n_video=600, load=0.7, check_every=27):
    """A 600-frame RGB8 1080p chrono-video (60-frame windows, stride 1 -> 540 windows over
    device-resident frames) interleaved with single-image jobs of 16-128 frames at 720p/1080p
    uploaded from pinned host memory.

    * plain: one request at a time (cp_outlier, image copied back), each waited for;
    * saturated: every request queued at once on cp_server (continuous batching, HBM-budget
      admission): its throughput against plain is the batching gain;
    * arrivals: the same requests submitted in wall time at seeded Poisson arrival times whose
      offered pixel-sample rate is `load` x the saturated throughput; latency (submit to
      result in host memory) is reported per class, windows and jobs;
    * results: every `check_every`-th window and every job, in both served runs, must equal the
      plain run's bytes."""
    import numpy as np
    rng = np.random.default_rng(seed)
    vsc = cpa.Scene(seed=seed, width=1920, height=1080, n_frames=n_video, n_disks=3, disk_r_min=10,
                    disk_r_max=40, blob_permille=100)
    video = _stack_from_scene(cpa, ctx, vsc)
    wins = [ix for _, ix in cpa.video_windows(cpa.FrameRange(0, window, 1),
                                               cpa.FrameRange(0, n_video - window, 1), n_video)]
    proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), cpa.BackgroundMode.FIRST,
                                cpa.OutlierSelectionMode.EXTREME)
    # host frame pools for the single-image jobs (pinned, copied from device-generated frames)
    pools = {}
    for (w, h) in ((1280, 720), (1920, 1080)):
        sc = cpa.Scene(seed=seed + w, width=w, height=h, n_frames=128, n_disks=2, disk_r_min=10,
                       disk_r_max=30, blob_permille=100)
        st = _stack_from_scene(cpa, ctx, sc)
        pools[(w, h)] = H.PinnedArray((128, h, w, 3), np.uint8)
        pools[(w, h)].array[...] = st.download()
        st.close()
    jobs_spec = []
    t = 0.0
    for i in range(n_jobs):
        t += rng.exponential(1.0 / n_jobs)  # position of the job among the windows
        n = int(rng.integers(16, 129))
        wh = (1280, 720) if rng.random() < 0.5 else (1920, 1080)
        jobs_spec.append((t, n, wh))
    # the request sequence: windows in order, each job after the window at its position
    reqs, ji = [], 0
    for k, ix in enumerate(wins):
        reqs.append(("window", k))
        while ji < len(jobs_spec) and jobs_spec[ji][0] * len(wins) <= k:
            reqs.append(("job", ji))
            ji += 1
    samples = {("window", k): len(ix) * 1920 * 1080 for k, ix in enumerate(wins)}
    samples.update({("job", i): n * wh[0] * wh[1] for i, (_, n, wh) in enumerate(jobs_spec)})
    total = sum(samples[r] for r in reqs)
    checked = [r for r in reqs if r[0] == "job" or r[1] % check_every == 0]
    outs_w = [H.PinnedArray((1080, 1920, 3), np.uint8) for _ in range(4)]  # unchecked windows rotate
    own_out = {r: H.PinnedArray((1080, 1920, 3) if r[0] == "window" else
                                (jobs_spec[r[1]][2][1], jobs_spec[r[1]][2][0], 3), np.uint8) for r in checked}

    # ---- plain: one request at a time, image back on the host
    plain_ref = {}
    t0 = time.perf_counter()
    for r in reqs:
        if r[0] == "window":
            img, _ = proc.process(video, image_indices=wins[r[1]], with_mask=False)
        else:
            _, n, (w, h) = jobs_spec[r[1]]
            st = cpa.Stack.from_array(ctx, pools[(w, h)].array[:n])
            img, _ = proc.process(st, with_mask=False)
            st.close()
        if r in own_out:
            plain_ref[r] = img
    plain_s = time.perf_counter() - t0

    def submit(srv, r):
        out = own_out[r].array if r in own_out else outs_w[r[1] % 4].array
        if r[0] == "window":
            job = cpa.Job(cpa.CP_MODE_OUTLIER, video, None, processor=proc, image_indices=wins[r[1]])
            return srv.submit(job, host_out=out)
        _, n, (w, h) = jobs_spec[r[1]]
        job = cpa.Job(cpa.CP_MODE_OUTLIER, None, None, processor=proc, n_frames=n, bytes_per_sample=1)
        return srv.submit(job, host_frames=pools[(w, h)].array[:n], host_out=out)

    def served(arrivals=None):
        for a in own_out.values():
            a.array[...] = 0
        srv = cpa.Server(ctx, hbm_budget=16 << 30, max_inflight=8)
        tickets = []
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            if arrivals is not None:
                while time.perf_counter() - t0 < arrivals[i]:
                    pass
            tickets.append(submit(srv, r))
        lat = np.array([srv.wait(tk) for tk in tickets])
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        match = all(np.array_equal(own_out[r].array, plain_ref[r]) for r in checked)
        return lat, wall, st, match

    # ---- saturated: everything queued at once
    lat_sat, sat_s, st_sat, match_sat = served()
    # ---- Poisson arrivals at `load` x the saturated throughput
    rate = load * total / sat_s  # offered pixel-samples per second
    gaps = rng.exponential(1.0, len(reqs)) * np.array([samples[r] for r in reqs]) / rate
    arrivals = np.cumsum(gaps) - gaps[0]
    lat, arr_s, st_arr, match_arr = served(arrivals)
    kinds = np.array([r[0] for r in reqs])

    def pct(a, q):
        return float(np.percentile(a, q)) if len(a) else None

    video.close()
    for a in list(own_out.values()) + outs_w + list(pools.values()):
        a.free()
    return {"windows": len(wins), "window_frames": window, "jobs": len(jobs_spec), "pixel_samples": int(total),
            "plain_value": total / plain_s, "served_value": total / sat_s, "speedup": plain_s / sat_s,
            "saturated": {"p50_ms": pct(lat_sat, 50), "p99_ms": pct(lat_sat, 99),
                          "peak_inflight": st_sat["peak_inflight"], "peak_hbm_bytes": st_sat["peak_bytes"],
                          "kernel_downloads": st_sat["kernel_downloads"]},
            "arrivals": {"process": "Poisson, seeded; mean gap = request pixel-samples / offered rate",
                         "offered_load": load, "offered_value": rate, "achieved_value": total / arr_s,
                         "window_p50_ms": pct(lat[kinds == "window"], 50),
                         "window_p99_ms": pct(lat[kinds == "window"], 99),
                         "job_p50_ms": pct(lat[kinds == "job"], 50), "job_p99_ms": pct(lat[kinds == "job"], 99),
                         "peak_inflight": st_arr["peak_inflight"]},
            "results_checked": len(checked) * 2, "results_match": bool(match_sat and match_arr),
            "plain": "one request at a time (cp_outlier, image copied to the host)",
            "served": "cp_server: 8 stream slots, 16 GiB budget, continuous admission, uploads on their own stream"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="outlier_4k_rgb16_256f", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-rows", type=int, default=1024, help="output rows of the CPU baseline band")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--force-generic", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batching and serving measurements")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale factor for HBM byte counters)")
    ap.add_argument("--batching", choices=sorted(BATCH_SETS),
                    help="only run the autotuner and this batching set (one JSON line), e.g. for a trace")
    args = ap.parse_args()
    if args.batching:
        import chrono_photo_amd as cpa
        from chrono_photo_amd import hip as H
        ctx = cpa.Context(0)
        cpa.autotune(ctx, 1, 3)
        print(json.dumps(run_batching(cpa, H, ctx, args.batching)))
        return
    if args.copy_calibration:
        from chrono_photo_amd import hip as H
        src, dst = H.DeviceBuffer(1 << 30), H.DeviceBuffer(1 << 30)
        for _ in range(3):
            H.check(H.hip().hipMemcpy(dst.ptr, src.ptr, 1 << 30, H.hipMemcpyDeviceToDevice))
        H.check(H.hip().hipDeviceSynchronize())
        print(json.dumps({"copy_calibration_bytes_read": 1 << 30, "copy_calibration_bytes_written": 1 << 30,
                          "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build composites on a single MI355X and has no "
              f"multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import chrono_photo_amd as cpa
    from chrono_photo_amd import hip as H

    wl = WORKLOADS[args.workload]
    scene = cpa.Scene(**wl["scene"])
    n, w, h, ch, bps = scene.n_frames, scene.width, scene.height, scene.channels, scene.bytes_per_sample
    ctx = cpa.Context(0)
    t0 = time.time()
    stack = cpa.Stack(ctx, w, h, ch, bps, capacity=n)
    scene.generate_device(stack.device_ptr, 0, n, stream=ctx.stream)
    ctx.synchronize()
    lib = ctx.lib
    cpa._check(lib.cp_stack_commit(stack.handle, n))
    gen_s = time.time() - t0
    print(f"[bench] {args.workload}: generated {n} frames {w}x{h}x{ch} u{8 * bps} in HBM in {gen_s:.1f}s",
          file=sys.stderr, flush=True)

    extra = {}
    crop = None
    if wl.get("shake"):
        sh = wl["shake"]
        t0 = time.time()
        offsets = cpa.ShakeAnalyzer().analyze(stack, sh["anchors"], sh["r"], sh["s"])
        shake_s = time.time() - t0
        truth = scene.shake_truth()
        crop = cpa.Crop.create(offsets, w, h)
        terms = len(sh["anchors"]) * (2 * sh["s"] + 1) ** 2 * (2 * sh["r"] + 1) ** 2 * ch * (n - 1)
        extra["shake"] = {"seconds": round(shake_s, 4), "sq_diff_terms_per_s": terms / shake_s,
                          "offsets_match_seed": bool(np.array_equal(offsets, truth)),
                          "frames_off_seed": int((offsets != truth).any(axis=1).sum()),
                          "anchors": len(sh["anchors"]), "radius": sh["r"], "search": sh["s"]}
        print(f"[bench] shake analysis {shake_s:.3f}s, offsets match seed: "
              f"{extra['shake']['offsets_match_seed']}", file=sys.stderr, flush=True)
    sel = cpa._FrameSel(stack, None, crop, n)
    oh, ow, _ = sel.out_shape(stack)

    out_px = H.DeviceBuffer(oh * ow * ch * bps)
    out_mask = H.DeviceBuffer(oh * ow * ch)
    flags = cpa.CP_FLAG_DEVICE_OUT | (cpa.CP_FLAG_FORCE_GENERIC if args.force_generic else 0)
    if wl["mode"] == "outlier":
        kind, lo, hi = wl["threshold"]
        thr = cpa.Threshold(kind == "abs", lo, hi)
        sample_pos = None
        if wl.get("sample"):
            # `--sample k` (src/chrono.rs:151-163): k distinct frame positions, sorted, seeded here
            sample_pos = np.sort(np.random.default_rng(7).choice(n, wl["sample"], replace=False))
        proc = cpa.OutlierProcessor(thr, cpa.BackgroundMode.from_str(wl["background"]),
                                    cpa.OutlierSelectionMode.from_str(wl["outlier"]),
                                    sample_indices=sample_pos)
        params = proc.params(bps)
        outs = cpa.cp_outputs(out_px.ptr, out_mask.ptr, None, None)

        def step():
            cpa._check(lib.cp_outlier(ctx.handle, stack.handle, C.byref(sel.c), C.byref(params),
                                      C.byref(outs), flags))
        out_bytes = oh * ow * ch * bps + oh * ow * ch
    else:
        weights = np.ones(4, np.float32)
        fade = cpa.Fade.none().native()

        def step():
            cpa._check(lib.cp_simple(ctx.handle, stack.handle, C.byref(sel.c),
                                     weights.ctypes.data_as(C.POINTER(C.c_float)), C.byref(fade),
                                     int(wl["mode"] == "darker"), C.c_void_p(out_px.ptr), flags))
        out_bytes = oh * ow * ch * bps

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    ev0, ev1 = H.Event(), H.Event()
    ev0.record(ctx.stream)
    for _ in range(args.steps):
        step()
    ev1.record(ctx.stream)
    ev1.synchronize()
    ctx.synchronize()
    ms = ev0.elapsed_ms(ev1) / args.steps
    pixel_samples = n * oh * ow
    value = pixel_samples / (ms / 1e3)
    alg_bytes = n * oh * ow * ch * bps + out_bytes
    achieved_gbs = alg_bytes / (ms / 1e3) / 1e9
    print(f"[bench] {ms:.3f} ms/step, {value:.4g} pixel-samples/s, {achieved_gbs:.0f} GB/s "
          f"({100 * achieved_gbs / HBM_PEAK_GBS:.1f}% of HBM peak)", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        rows = min(args.cpu_rows, oh)
        # band of the stack that covers output rows [0, rows) under every frame's crop
        band_h = rows + (int(crop[0][:, 1].max()) if crop is not None else 0)
        band = np.zeros((n, band_h, w, ch), np.uint16 if bps == 2 else np.uint8)
        H.copy_rows_to_host(band, stack.device_ptr, w * h * ch * bps, n)
        threads = os.cpu_count() or 1
        if wl["mode"] == "outlier":
            othr = O.threshold(thr.absolute, thr.min, thr.max, bps)
            bgm, selm = int(proc.background), int(proc.outlier)
            kw = dict(crop_xy=None if crop is None else crop[0], out_w=ow, out_h=rows, sample_pos=sample_pos)
            t0 = time.time()
            ref_img, _, _ = O.outlier(band, othr, bgm, selm, threads=threads, y0=0, y1=rows, **kw)
            cpu_s = time.time() - t0
            one_rows = max(1, rows // 16)
            t0 = time.time()
            O.outlier(band, othr, bgm, selm, threads=1, y0=0, y1=one_rows, **kw)
            one_s = time.time() - t0
        else:
            kw = dict(out_w=ow, out_h=rows)
            t0 = time.time()
            ref_img = O.simple(band, darker=wl["mode"] == "darker", threads=threads, y0=0, y1=rows, **kw)
            cpu_s = time.time() - t0
            one_rows = max(1, rows // 16)
            t0 = time.time()
            O.simple(band, darker=wl["mode"] == "darker", threads=1, y0=0, y1=one_rows, **kw)
            one_s = time.time() - t0
        gpu_img = out_px.to_host(np.zeros((rows, ow, ch), np.uint16 if bps == 2 else np.uint8))
        match = bool(np.array_equal(gpu_img, ref_img[:rows]))
        cpu = {"value": n * rows * ow / cpu_s, "unit": "pixel-samples/s", "cores": threads, "kind": "port",
               "sample": f"output rows [0,{rows}) x {ow} px x {n} frames (CPU oracle, OpenMP row-parallel)",
               "single_thread_value": n * one_rows * ow / one_s, "gpu_rows_bit_exact": match}
        print(f"[bench] cpu oracle {cpu['value']:.4g} px-samples/s on {threads} threads, "
              f"{cpu['single_thread_value']:.4g} on 1 thread; GPU rows bit-exact: {match}", file=sys.stderr)

    if not args.no_extras and args.workload == "outlier_4k_rgb16_256f":
        stack.close()
        out_px.free()
        out_mask.free()
        # the 8-bit workloads: 64-frame relative-threshold average, and the tall galleries
        extra["workloads"] = {}
        for name in ("outlier_1080p_rgb8_64f", "outlier_1080p_rgb8_1200f", "outlier_1080p_rgb8_1200f_s64"):
            t0 = time.time()
            r = extra["workloads"][name] = run_workload(cpa, H, ctx, name)
            print(f"[bench] {name}: {r['ms_per_step']:.3f} ms/step, {r['value']:.4g} pixel-samples/s, "
                  f"{r['roofline']['achieved']:.0f} GB/s ({100 * r['roofline']['frac']:.1f}% of HBM peak) "
                  f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        t0 = time.time()
        tuned = cpa.autotune(ctx, 1, 3)  # the batching sets are 8-bit RGB
        extra["autotune_u8_rgb"] = {kind: [None if sh is None else "x".join(map(str, sh)) for sh in v]
                                    for kind, v in tuned.items()}
        print(f"[bench] autotune (8-bit RGB, G x F x X x kernel per bucket): {extra['autotune_u8_rgb']} "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        t0 = time.time()
        extra["batching"] = run_batching(cpa, H, ctx, "large")
        bt = extra["batching"]
        print(f"[bench] batching (48 large stacks): {bt['speedup']:.2f}x over plain, "
              f"{bt['plain_ms']:.2f} -> {bt['batched_ms']:.2f} ms (queued {bt['queued_ms']:.2f} ms); "
              f"launches {bt['launches']}; byte-identical {bt['byte_identical']} "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        t0 = time.time()
        extra["batching_small"] = run_batching(cpa, H, ctx, "small")
        bt = extra["batching_small"]
        print(f"[bench] batching (256 small stacks): {bt['speedup']:.2f}x over plain, "
              f"{bt['plain_ms']:.2f} -> {bt['batched_ms']:.2f} ms (queued {bt['queued_ms']:.2f} ms); "
              f"launches {bt['launches']}; byte-identical {bt['byte_identical']} "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        t0 = time.time()
        extra["serving"] = sv = run_serving(cpa, H, ctx)
        print(f"[bench] serving: {sv['speedup']:.2f}x over plain saturated; Poisson arrivals at "
              f"{sv['arrivals']['offered_load']:.0%} load: windows p50 {sv['arrivals']['window_p50_ms']:.2f} ms "
              f"p99 {sv['arrivals']['window_p99_ms']:.2f} ms, jobs p50 {sv['arrivals']['job_p50_ms']:.2f} ms "
              f"p99 {sv['arrivals']['job_p99_ms']:.2f} ms; results match: {sv['results_match']} "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    traffic = measured_traffic(args.workload)
    res = {
        "metric": "pixel-samples/s outlier mode (% HBM roofline)" if wl["mode"] == "outlier"
        else f"pixel-samples/s {wl['mode']} mode (% HBM roofline)",
        "value": value, "unit": "pixel-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16" if bps == 2 else "u8", "data": "synthetic",
        "config": {"workload": args.workload, "frames": n, "width": w, "height": h, "channels": ch,
                   "out_width": ow, "out_height": oh, "mode": wl["mode"],
                   **({"threshold": "/".join(map(str, wl["threshold"])), "background": wl["background"],
                       "outlier": wl["outlier"]} if wl["mode"] == "outlier" else {}),
                   **({"sample": wl["sample"]} if wl.get("sample") else {}),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     **({"traffic_source": traffic["source"]} if traffic else {}),
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(res))
    ctx.close()


if __name__ == "__main__":
    main()
