"""GPU parity: the HIP kernels behind the C ABI against the CPU oracle, bit-exact.

Integer outputs (pixels, blend masks, outlier counts, selected frames) and the f32 order
statistics (medians, inverse IQR) must match exactly.  Sizes are small so the oracle
finishes in well under a second per case."""
import itertools

import numpy as np
import pytest

import chrono_photo_amd as cpa
import oracle_py as O

pytestmark = pytest.mark.gpu


def make_stack(seed, n, h, w, c, bits=8, outlier_frac=0.15):
    rng = np.random.default_rng(seed)
    hi = (1 << bits) - 1
    base = rng.integers(hi // 4, 3 * hi // 4, (1, h, w, c))
    noise = rng.integers(-(hi // 40) - 1, hi // 40 + 2, (n, h, w, c))
    x = np.clip(base + noise, 0, hi)
    hit = rng.random((n, h, w)) < outlier_frac
    x = np.where(hit[..., None], rng.integers(0, hi + 1, (n, h, w, c)), x)
    # a few all-outlier and tied pixels
    x[:, 0, 0, :] = np.where(np.arange(n)[:, None] % 2 == 0, 0, hi)
    x[:, 0, 1, :] = 7
    return x.astype(np.uint8 if bits == 8 else np.uint16)


def run_both(ctx, x, thr_abs, lo, hi, bg, sel, weights=(1, 1, 1, 1), fade=None, slots=None,
             crop=None, sample=None, seed=42, force_generic=False):
    bps = x.dtype.itemsize
    st = cpa.Stack.from_array(ctx, x)
    thr = cpa.Threshold(thr_abs, lo, hi)
    cfade = cpa.Fade.none() if fade is None else cpa.Fade(fade[0], fade[1], fade[2])
    proc = cpa.OutlierProcessor(thr, bg, sel, weights, cfade, sample_indices=sample, seed=seed)
    cr = None
    if crop is not None:
        cr = crop
    img, mask, stats = proc.process(st, image_indices=slots, crop=cr, with_stats=True,
                                    force_generic=force_generic)
    ofade = O.Fade.none() if fade is None else O.Fade(fade[0], fade[1], [f for f, _ in fade[2]],
                                                        [v for _, v in fade[2]])
    oimg, omask, owarn, ost = O.outlier(
        x, O.threshold(thr_abs, lo, hi, bps), bg, sel, weights=weights, fade=ofade, seed=seed,
        slots=slots, crop_xy=None if crop is None else crop[0],
        out_w=None if crop is None else crop[1], out_h=None if crop is None else crop[2],
        sample_pos=sample, stats=True)
    return (img, mask, stats, proc.warnings), (oimg, omask, ost, owarn)


def assert_same(gpu, ref, ctx_msg=""):
    img, mask, st, warn = gpu
    oimg, omask, ost, owarn = ref
    assert np.array_equal(st["median"], ost["median"]), ctx_msg + " median"
    assert np.array_equal(st["iqr_inv"], ost["iqr_inv"]), ctx_msg + " iqr_inv"
    assert np.array_equal(st["num_outliers"], ost["num_outliers"]), ctx_msg + " num_outliers"
    assert np.array_equal(st["max_index"], ost["max_index"]), ctx_msg + " max_index"
    assert np.array_equal(st["max_dist_sq"], ost["max_dist_sq"]), ctx_msg + " max_dist_sq"
    assert np.array_equal(st["bg_index"], ost["bg_index"]), ctx_msg + " bg_index"
    assert np.array_equal(st["warning"], ost["warning"]), ctx_msg + " warning"
    bad = np.argwhere(img != oimg)
    assert bad.size == 0, f"{ctx_msg} pixels differ at {bad[:5].tolist()} gpu={img[tuple(bad[0])]} ref={oimg[tuple(bad[0])]}"
    assert np.array_equal(mask, omask), ctx_msg + " mask"
    assert warn == owarn, ctx_msg + " warnings"


@pytest.mark.parametrize("c,w,shift", [(3, 20, 0), (3, 21, 0), (4, 20, 0), (4, 21, 0), (3, 20, 2), (4, 21, 2)])
def test_outlier_u16_wrapped_stack(ctx, c, w, shift):
    """A borrowed allocation ends at its last frame and may start off an 8-byte boundary: the
    tiles that read a pixel pair per load fall back to one pixel per lane where a pair read
    would pass that end or start misaligned."""
    from chrono_photo_amd import hip as H
    n, h = 40, 9
    x = np.ascontiguousarray(make_stack(900 + 10 * c + w, n, h, w, c, bits=16))
    buf = H.DeviceBuffer(x.nbytes + shift)
    H.check(H.hip().hipMemcpy(buf.ptr + shift, x.ctypes.data, x.nbytes, H.hipMemcpyHostToDevice), "hipMemcpy")
    st = cpa.Stack(ctx, w, h, c, 2, device_ptr=buf.ptr + shift, n_frames=n)
    for sel in (2, 4):
        proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), 0, sel)
        img, mask = proc.process(st)
        oimg, omask, _ = O.outlier(x, O.threshold(True, 0.05, 0.2, 2), 0, sel)
        assert np.array_equal(img, oimg), (c, w, sel)
        assert np.array_equal(mask, omask), (c, w, sel)
    st.close()
    buf.free()


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
@pytest.mark.parametrize("c", [3, 4])
def test_outlier_u16_full_tiles(ctx, n, c):
    """Frame counts that fill the tile exactly: absolute thresholds take the byte search in a
    128-wide window around the remedian estimate (the window next to it, then the exact
    two-level search, where the result lands on a window's edge), relative ones the exact
    two-level search for the median and both quartiles; noisy pixels, pixels with runs of
    outliers, constant pixels and pixels at either end of the range all stay exact."""
    for frac, seed in ((0.02, 1), (0.3, 2)):
        x = make_stack(1300 + 10 * n + c + seed, n, 5, 70, c, bits=16, outlier_frac=frac)
        x[:, 1, 2:6, :] = 40000                        # constant pixels
        x[: n // 3, 2, 2:6, :] = 1000                  # a run of outliers in the first blocks
        x[:, 3, 2:6, :] = 65535 - (np.arange(n) % 3)[:, None, None]  # at the top of the range
        x[:, 4, 2:6, :] = np.arange(n)[:, None, None] % 2            # at the bottom
        for sel in (2, 4):
            gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 0, sel)
            assert_same(gpu, ref, f"n={n} c={c} frac={frac} sel={sel}")
        gpu, ref = run_both(ctx, x, False, 2.0, 4.0, 2, 3)
        assert_same(gpu, ref, f"n={n} c={c} frac={frac} rel")


def _window_offsets(x):
    """Distance of each 16-bit median from the estimate that places its search window, for
    lanes of 32 frames: per lane the median of three medians of three over the half-words of
    its first nine frame pairs, averaged over both halves, then the mean over the lanes
    without the smallest and the largest (the plain mean for two lanes)."""
    n = x.shape[0]
    g = n // 32
    v = np.moveaxis(x.astype(np.int64), 0, -1).reshape(x.shape[1:] + (g, 32))
    m3 = lambda a, b, c: np.maximum(np.minimum(a, b), np.minimum(np.maximum(a, b), c))
    e = 0
    for h in (0, 1):
        d = [v[..., 2 * i + h] for i in range(9)]
        e = e + m3(m3(d[0], d[1], d[2]), m3(d[3], d[4], d[5]), m3(d[6], d[7], d[8]))
    e = e >> 1
    est = e.sum(-1) >> 1 if g == 2 else (e.sum(-1) - e.min(-1) - e.max(-1)) // (g - 2)
    med = np.sort(v.reshape(v.shape[:-2] + (n,)), axis=-1)[..., (n + 1) // 2 - 1]
    return med - est


@pytest.mark.parametrize("n", [64, 128, 256])
def test_outlier_u16_median_windows(ctx, n):
    """Absolute-threshold medians of 16-bit stacks under every candidate tile shape, with the
    estimate pushed away from the median pixel by pixel: 12 of each 32-frame lane block (two
    of each triple of the frame pairs the estimate reads) sit D above or below the rest, so
    the estimate lands near them while the median stays with the majority.  D from 0 to 945
    puts the median inside the first 128-wide window, inside the window next to it, and
    outside both (the exact two-level search); each occurs, and every result is exact."""
    c, h, w = 3, 2, 136
    rng = np.random.default_rng(4100 + n)
    base = rng.integers(20000, 40000, (1, h, w, c))
    x = base + rng.integers(-4, 5, (n, h, w, c))
    j = np.arange(n) % 32
    shifted = np.isin(j, [0, 1, 2, 3, 6, 7, 8, 9, 12, 13, 14, 15])
    d = (np.arange(w) * 7)[None, None, :, None] * np.array([1, -1, 1])
    d = np.where(np.arange(h)[None, :, None, None] == 1, d // 2 + 3, d)
    x = np.where(shifted[:, None, None, None], x + d, x).astype(np.uint16)
    off = np.abs(_window_offsets(x))
    if n == 256:
        assert (off <= 60).any() and ((off >= 70) & (off <= 180)).any() and (off >= 200).any()
    try:
        for pick in PICKS:
            shapes = [cands[min(pick, len(cands) - 1)] for cands in CANDIDATES[16]]
            cpa.tile_shapes_set(ctx, 2, c, shapes)
            for sel in (0, 4):
                gpu, ref = run_both(ctx, x, True, 0.004, 0.02, 0, sel)
                assert_same(gpu, ref, f"n={n} shape={shapes} sel={sel}")
    finally:
        cpa.tile_shapes_set(ctx, 2, c, [])


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
@pytest.mark.parametrize("c", [3, 4])
def test_outlier_u8_full_tiles(ctx, n, c):
    """8-bit stacks whose frame count fills a tile shape exactly, with constant pixels, a run
    of outliers and pixels at either end of the range, for absolute and relative thresholds."""
    x = make_stack(1700 + n + c, n, 5, 40, c, bits=8, outlier_frac=0.1)
    x[:, 1, 2:6, :] = 200
    x[: n // 3, 2, 2:6, :] = 10
    x[:, 3, 2:6, :] = 255 - (np.arange(n) % 3)[:, None, None]
    x[:, 4, 2:6, :] = np.arange(n)[:, None, None] % 2
    for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3), (False, 1.0, 1.5, 3, 5)):
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel)
        assert_same(gpu, ref, f"n={n} c={c} abs={absolute} sel={sel}")


def oracle_like(proc, x, slots=None, crop=None):
    """The oracle's composite for an OutlierProcessor's options: (image, mask, stats, warnings)."""
    bps = x.dtype.itemsize
    t = proc.threshold
    fade = O.Fade.none()
    if not proc.fade.is_none:
        raise ValueError("oracle_like covers composites without fade")
    img, mask, warn, st = O.outlier(
        x, O.threshold(t.absolute, t.min, t.max, bps), int(proc.background), int(proc.outlier),
        weights=tuple(float(v) for v in proc.weights), fade=fade, seed=proc.seed, slots=slots,
        crop_xy=None if crop is None else crop[0], out_w=None if crop is None else crop[1],
        out_h=None if crop is None else crop[2], sample_pos=proc.sample_indices, stats=True)
    return img, mask, st, warn


def assert_host_matches_oracle(got, warnings, ref, msg):
    """An out-of-core result (image, mask, stats) and its warning count against the oracle."""
    img, mask, st = got
    assert_same((img, mask, st, warnings), ref, msg)


@pytest.mark.parametrize("bits,c", [(8, 3), (16, 3), (16, 4), (8, 1)])
def test_outlier_host_bands_equal_resident(ctx, bits, c):
    """Out-of-core composites (row bands of host frames streamed through a band stack) equal
    the oracle and the resident composite: backgrounds incl. random, outlier modes, crops, a
    frame selection, statistics and warnings, for bands of 1, 3 and all rows."""
    n, h, w = 40, 23, 37
    x = make_stack(1500 + bits + c, n, h, w, c, bits=bits)
    rng = np.random.default_rng(5)
    crop_xy = rng.integers(0, 5, (n, 2)).astype(np.int32)
    crop = (crop_xy, w - 4, h - 4)
    st = cpa.Stack.from_array(ctx, x)
    row = w * c * (bits // 8)
    for bg, sel in ((1, 4), (0, 2), (2, 3), (3, 5)):
        proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), bg, sel, seed=9)
        ref = proc.process(st, crop=crop, with_stats=True)
        ref_warn = proc.warnings
        oref = oracle_like(proc, x, crop=crop)
        for band_rows in (1, 3, 64):
            band_bytes = (band_rows + 4) * row * n + 256  # crop halo of at most 4 rows + slack
            got = proc.process_host(ctx, x, crop=crop, with_stats=True, band_bytes=band_bytes)
            assert_host_matches_oracle(got, proc.warnings, oref, f"bits={bits} c={c} bg={bg} sel={sel} band={band_rows}")
            for a, b, what in zip(got, ref, ("image", "mask", "stats")):
                assert np.array_equal(a, b), (bits, c, bg, sel, band_rows, what)
            assert proc.warnings == ref_warn
    # a frame selection (the reference's image_indices) without crops
    idx = [3, 4, 9, 10, 11, 20, 31, 32, 38]
    proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), 0, 0)
    ref = proc.process(st, image_indices=idx)
    got = proc.process_host(ctx, x, image_indices=idx, band_bytes=2 * row * len(idx) + 256, with_stats=True)
    assert_host_matches_oracle(got, proc.warnings, oracle_like(proc, x, slots=idx), "frame selection")
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    with pytest.raises(cpa.ChronoError) as e:
        proc.process_host(ctx, x, crop=crop, band_bytes=row * n)
    assert e.value.code == cpa.CP_ERR_OUT_OF_HBM_BUDGET
    st.close()


def test_host_bands_random_configurations(ctx):
    """40 out-of-core composites from a fixed seed (band heights from 1 row to the whole
    image, crops, frame selections, outlier and darker/lighter modes, 1-4 channels) equal the
    oracle and the resident composite, statistics included."""
    rng = np.random.default_rng(20261206)
    for case in range(40):
        bits = int(rng.choice([8, 16]))
        c = int(rng.integers(1, 5))
        n = int(rng.integers(3, 60))
        h, w = int(rng.integers(6, 24)), int(rng.integers(6, 40))
        x = make_stack(int(rng.integers(1 << 30)), n, h, w, c, bits=bits, outlier_frac=float(rng.uniform(0, 0.4)))
        st = cpa.Stack.from_array(ctx, x)
        slots = None
        if rng.random() < 0.3:
            slots = sorted(rng.choice(n, int(rng.integers(3, n + 1)), replace=False).tolist())
        m = n if slots is None else len(slots)
        crop, halo = None, 0
        if rng.random() < 0.5:
            offs = rng.integers(-2, 3, (m, 2)).tolist()
            offs[0] = [0, 0]
            cr = O.crop_create(offs, w, h)
            if cr is not None and cr[1] > 0 and cr[2] > 0:
                crop = cr
                halo = int(cr[0][:, 1].max() - cr[0][:, 1].min())
        row = w * c * (bits // 8)
        band_rows = int(rng.integers(1, h + 3))
        band_bytes = (band_rows + halo) * row * m + 256
        if rng.random() < 0.7:
            absolute = bool(rng.random() < 0.5)
            thr = cpa.Threshold.abs(0.05, 0.2) if absolute else cpa.Threshold.rel(2.0, 4.0)
            proc = cpa.OutlierProcessor(thr, int(rng.integers(0, 4)), int(rng.integers(0, 6)),
                                        seed=int(rng.integers(1 << 31)))
            ref = proc.process(st, image_indices=slots, crop=crop, with_stats=True)
            ref_warn = proc.warnings
            got = proc.process_host(ctx, x, image_indices=slots, crop=crop, with_stats=True, band_bytes=band_bytes)
            assert_host_matches_oracle(got, proc.warnings, oracle_like(proc, x, slots=slots, crop=crop),
                                       f"case {case}: bits={bits} c={c} n={n} {w}x{h} band={band_rows}")
            for a, b, what in zip(got, ref, ("image", "mask", "stats")):
                assert np.array_equal(a, b), f"case {case}: {what} bits={bits} c={c} n={n} {w}x{h} band={band_rows}"
            assert proc.warnings == ref_warn, f"case {case}: warnings"
        else:
            proc = cpa.SimpleProcessor(darker=bool(rng.random() < 0.5))
            ref = proc.process(st, image_indices=slots, crop=crop)
            got = proc.process_host(ctx, x, image_indices=slots, crop=crop, band_bytes=band_bytes)
            oimg = O.simple(x, darker=proc.darker, slots=slots, crop_xy=None if crop is None else crop[0],
                            out_w=None if crop is None else crop[1], out_h=None if crop is None else crop[2])
            assert np.array_equal(got, oimg), f"case {case}: simple vs oracle bits={bits} c={c} n={n} {w}x{h}"
            assert np.array_equal(got, ref), f"case {case}: simple bits={bits} c={c} n={n} {w}x{h} band={band_rows}"
        st.close()


@pytest.mark.parametrize("bits,c", [(8, 3), (16, 4)])
def test_simple_host_bands_equal_resident(ctx, bits, c):
    """Darker / lighter over host frames in row bands equal the oracle and the resident
    composite, with crops, a frame selection and a fade."""
    n, h, w = 30, 21, 33
    x = make_stack(1600 + bits + c, n, h, w, c, bits=bits)
    rng = np.random.default_rng(6)
    crop = (rng.integers(0, 4, (n, 2)).astype(np.int32), w - 3, h - 3)
    st = cpa.Stack.from_array(ctx, x)
    row = w * c * (bits // 8)
    keys = [(0, 0.2), (12, 1.0), (25, 0.5)]
    fade = cpa.Fade(False, True, keys)
    ofade = O.Fade(False, True, [f for f, _ in keys], [v for _, v in keys])
    for darker in (True, False):
        proc = cpa.SimpleProcessor(fade=fade, darker=darker)
        ref = proc.process(st, crop=crop)
        oimg = O.simple(x, fade=ofade, darker=darker, crop_xy=crop[0], out_w=crop[1], out_h=crop[2])
        for band_rows in (1, 5, 64):
            got = proc.process_host(ctx, x, crop=crop, band_bytes=(band_rows + 3) * row * n + 256)
            assert np.array_equal(got, oimg), (bits, c, darker, band_rows, "oracle")
            assert np.array_equal(got, ref), (bits, c, darker, band_rows)
        idx = list(range(2, 28, 3))
        ref = proc.process(st, image_indices=idx)
        got = proc.process_host(ctx, x, image_indices=idx, band_bytes=4 * row * len(idx) + 256)
        assert np.array_equal(got, O.simple(x, fade=ofade, darker=darker, slots=idx)), (bits, c, darker, "oracle frames")
        assert np.array_equal(got, ref), (bits, c, darker, "frames")
    st.close()


MODES = list(itertools.product([0, 1, 2, 3], [0, 1, 2, 3, 4, 5]))


@pytest.mark.parametrize("bg,sel", MODES)
@pytest.mark.parametrize("absolute", [True, False])
def test_outlier_all_modes_u8_rgb(ctx, bg, sel, absolute):
    x = make_stack(100 + bg * 7 + sel, 17, 13, 19, 3)
    lo, hi = (0.05, 0.2) if absolute else (2.0, 4.0)
    for generic in (False, True):
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, force_generic=generic)
        assert_same(gpu, ref, f"bg={bg} sel={sel} abs={absolute} generic={generic}")


@pytest.mark.parametrize("sel", [0, 2, 3, 4, 5])
@pytest.mark.parametrize("absolute", [True, False])
def test_outlier_u16_rgb(ctx, sel, absolute):
    x = make_stack(300 + sel, 21, 11, 14, 3, bits=16)
    lo, hi = (0.05, 0.2) if absolute else (2.0, 4.0)
    gpu, ref = run_both(ctx, x, absolute, lo, hi, 0, sel)
    assert_same(gpu, ref, f"u16 sel={sel}")


@pytest.mark.parametrize("c", [1, 2, 4])
def test_outlier_channel_counts_and_weights(ctx, c):
    x = make_stack(500 + c, 15, 9, 10, c)
    for w in [(1, 1, 1, 1), (0.3, 0, 2.0, 1.0), (-1.0, 0.5, 1.0, 0.0)]:
        gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 2, 3, weights=w)
        assert_same(gpu, ref, f"c={c} w={w}")
        gpu, ref = run_both(ctx, x, False, 1.5, 3.0, 3, 4, weights=w)
        assert_same(gpu, ref, f"c={c} w={w} rel")


@pytest.mark.parametrize("bits,c,n", [(16, 3, 200), (8, 3, 120), (8, 4, 60), (16, 4, 30), (16, 3, 30),
                                      (16, 3, 100), (16, 4, 60), (16, 3, 12)])
def test_outlier_many_tiles_with_crops(ctx, bits, c, n):
    """Many tiles per row, a last tile per row that is partly past the image, and per-frame
    crop origins that put every frame row at a different alignment."""
    x = make_stack(1100 + bits + c + n, n, 40, 1003, c, bits=bits, outlier_frac=0.05)
    rng = np.random.default_rng(n)
    crop_xy = rng.integers(0, 7, (n, 2)).astype(np.int32)
    crop = (crop_xy, 996, 33)
    for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3)):
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, crop=crop)
        assert_same(gpu, ref, f"bits={bits} c={c} n={n} abs={absolute}")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 31, 32, 33, 64, 65, 100])
def test_outlier_frame_counts(ctx, n):
    x = make_stack(700 + n, n, 7, 9, 3)
    gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 0, 4)
    assert_same(gpu, ref, f"n={n}")
    if n >= 3:
        gpu, ref = run_both(ctx, x, False, 2.0, 4.0, 2, 3)
        assert_same(gpu, ref, f"n={n} rel")


# each frame bucket's candidate kernel choices (lanes per pixel, frames per lane, pixels per
# lane[, 1: the 8-bit packed-byte kernel]) for 3-4 channel stacks; buckets <= 16, 32, 64, 128,
# 256, 512, 768, 1024, 1280 frames
CANDIDATES = {
    8: [[(1, 16, 1), (1, 16, 1, 1), (1, 16, 1, 2)], [(1, 32, 1), (2, 16, 1), (1, 32, 1, 1), (1, 32, 1, 2)],
        [(2, 32, 1), (4, 16, 1), (2, 32, 1, 1), (1, 64, 1, 2), (1, 64, 1, 1)],
        [(4, 32, 1), (8, 16, 1), (4, 32, 1, 1)], [(8, 32, 1), (4, 64, 1), (8, 32, 1, 1)],
        [(8, 64, 1), (16, 32, 4, 1), (16, 32, 1, 1)],
        [(32, 24, 4, 1)], [(32, 32, 4, 1)], [(32, 40, 4, 1)]],
    16: [[(1, 16, 2), (1, 16, 1)], [(1, 32, 2), (1, 32, 1)], [(2, 32, 2), (2, 32, 1)], [(4, 32, 2), (4, 32, 1)],
         [(8, 32, 2), (8, 32, 1)], [(8, 64, 1)], [None], [None], [None]],
}
PICKS = (0, 1, 2, 3, 4)


@pytest.mark.parametrize("bits,c", [(8, 3), (8, 4), (16, 3), (16, 4)])
def test_outlier_every_candidate_tile_shape(ctx, bits, c):
    """Every tile shape the autotuner can pick, pinned with cp_tile_shapes_set: frame counts
    that fill each bucket, a run of outliers longer than one 16-frame block (the background
    is then the first frame after the run), constant and range-edge pixels."""
    hi = (1 << bits) - 1
    try:
        for pick in PICKS:
            shapes = [cands[min(pick, len(cands) - 1)] for cands in CANDIDATES[bits]]
            cpa.tile_shapes_set(ctx, bits // 8, c, shapes)
            for n in (16, 32, 64, 128, 256) + ((512, 768, 1024, 1280) if bits == 8 else ()):
                x = make_stack(2500 + n + c + bits, n, 4, 37, c, bits=bits, outlier_frac=0.05)
                x[:, 1, 2:6, :] = hi // 2
                x[: (n - 1) // 2, 2, 2:6, :] = hi // 40          # outliers over the first blocks
                x[:, 3, 2:6, :] = hi - (np.arange(n) % 3)[:, None, None]
                for absolute, lo, thr_hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3)):
                    gpu, ref = run_both(ctx, x, absolute, lo, thr_hi, bg, sel)
                    assert_same(gpu, ref, f"bits={bits} c={c} n={n} shape={shapes} abs={absolute}")
    finally:
        cpa.tile_shapes_set(ctx, bits // 8, c, [])


@pytest.mark.parametrize("bits", [8, 16])
def test_every_candidate_tile_shape_all_modes(ctx, bits):
    """All background x selection modes, absolute and relative, under every candidate choice
    of the 64- and 128-frame buckets."""
    hi = (1 << bits) - 1
    try:
        for pick in PICKS:
            shapes = [cands[min(pick, len(cands) - 1)] for cands in CANDIDATES[bits]]
            cpa.tile_shapes_set(ctx, bits // 8, 3, shapes)
            for n in (64, 128):
                x = make_stack(2700 + n + bits + pick, n, 3, 21, 3, bits=bits, outlier_frac=0.1)
                x[: (n - 1) // 2, 1, 2:6, :] = hi // 40
                for bg, sel in MODES:
                    for absolute, lo, thr_hi in ((True, 0.05, 0.2), (False, 2.0, 4.0)):
                        gpu, ref = run_both(ctx, x, absolute, lo, thr_hi, bg, sel)
                        assert_same(gpu, ref, f"bits={bits} n={n} pick={pick} bg={bg} sel={sel} abs={absolute}")
    finally:
        cpa.tile_shapes_set(ctx, bits // 8, 3, [])


@pytest.mark.parametrize("bits", [8, 16])
def test_finish_outlier_counts_every_shape(ctx, bits):
    """The finish re-reads the outliers' samples up to four at a time before it blends them
    in frame order (forward, backward) or sums their distances (average): columns with
    exactly k outliers (k = 0-5, 7-9, 16, 17, 23, at most a third of the stack), whole
    batches and partial last batches, at random frames and at distances between the two
    limits, so every blend is partial and the frame order shows in the image, under every
    candidate tile shape (the 600-frame 8-bit stack runs the four-pixel-per-lane tiles)."""
    hi = (1 << bits) - 1
    ks = [0, 1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 23]
    rng = np.random.default_rng(5100 + bits)
    stacks = []
    for n in (32, 64, 128) + ((600,) if bits == 8 else (256,)):
        h, w = 2, 4 * len(ks)
        x = np.full((n, h, w, 3), hi // 2, dtype=np.int64)
        x += rng.integers(-(hi // 100) - 1, hi // 100 + 2, x.shape)
        want = np.zeros((h, w), dtype=np.int64)
        for i, kk in enumerate(ks):
            kk = min(kk, n // 3)
            for col in range(4 * i, 4 * i + 4):
                frames = rng.choice(n, kk, replace=False)
                off = rng.uniform(0.06, 0.16, (kk, h, 1)) * hi * rng.choice([-1, 1], (kk, h, 1))
                x[frames, :, col] = np.clip(hi // 2 + off, 0, hi).astype(np.int64)
                want[:, col] = kk
        stacks.append((x.astype(np.uint8 if bits == 8 else np.uint16), want))
    try:
        for pick in PICKS:
            shapes = [cands[min(pick, len(cands) - 1)] for cands in CANDIDATES[bits]]
            cpa.tile_shapes_set(ctx, bits // 8, 3, shapes)
            for x, want in stacks:
                for bg, sel in ((0, 4), (0, 5), (0, 3), (3, 4), (2, 3)):
                    gpu, ref = run_both(ctx, x, True, 0.05, 0.2, bg, sel)
                    msg = f"bits={bits} n={x.shape[0]} pick={pick} bg={bg} sel={sel}"
                    assert_same(gpu, ref, msg)
                    assert np.array_equal(gpu[2]["num_outliers"].reshape(want.shape), want), msg
    finally:
        cpa.tile_shapes_set(ctx, bits // 8, 3, [])


def test_random_configurations(ctx):
    """300 configurations drawn from a fixed seed: sample width, 1-4 channels, 1-300 frames,
    image shape, threshold kind and limits, background and selection mode, weights, fade,
    frame selection, crop and pinned tile shape, each against the oracle."""
    rng = np.random.default_rng(20261203)
    for case in range(300):
        bits = int(rng.choice([8, 16]))
        c = int(rng.integers(1, 5))
        n = int(rng.choice([int(rng.integers(1, 40)), int(rng.integers(40, 300))]))
        w, h = int(rng.integers(8, 60)), int(rng.integers(4, 7))
        x = make_stack(int(rng.integers(1 << 30)), n, h, w, c, bits=bits, outlier_frac=float(rng.uniform(0, 0.5)))
        slots = None
        if n > 4 and rng.random() < 0.3:
            slots = sorted(rng.choice(n, int(rng.integers(3, n)), replace=False).tolist())
        m = n if slots is None else len(slots)
        absolute = m < 3 or rng.random() < 0.5
        lo = float(rng.uniform(0.01, 0.3)) if absolute else float(rng.uniform(0.5, 3.0))
        hi = lo + float(rng.uniform(0, 0.3 if absolute else 2.0))
        bg, sel = int(rng.integers(0, 4)), int(rng.integers(0, 6))
        weights = tuple(float(v) for v in rng.choice([1.0, 0.5, 2.0, 0.0, -1.0], 4))
        if all(v == 0.0 for v in weights[:c]):
            weights = (1.0,) + weights[1:]
        fade = None
        if rng.random() < 0.4:
            keys = sorted(rng.choice(max(n, 4), int(rng.integers(2, 5)), replace=False).tolist())
            fade = (bool(rng.random() < 0.5), bool(rng.random() < 0.5),
                    [(int(k), float(rng.choice([0.0, 0.25, 0.5, 1.0, float(rng.random())]))) for k in keys])
        crop = None
        if rng.random() < 0.3:
            offs = rng.integers(-2, 3, (m, 2)).tolist()
            offs[0] = [0, 0]
            cr = O.crop_create(offs, w, h)
            if cr is not None and cr[1] > 0 and cr[2] > 0:  # +-2 on 4 rows can leave nothing
                crop = cr
        sample = None
        if rng.random() < 0.25:  # `--sample`: medians over a sorted subset of the positions
            k = int(rng.integers(3 if not absolute else 1, m + 1))
            sample = sorted(rng.choice(m, k, replace=False).tolist())
        if c >= 3 and rng.random() < 0.5:
            cands = CANDIDATES[bits]
            cpa.tile_shapes_set(ctx, bits // 8, c, [cs[int(rng.integers(0, len(cs)))] for cs in cands])
        try:
            gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, weights=weights, fade=fade, slots=slots,
                                crop=crop, sample=sample, seed=int(rng.integers(1 << 31)))
            assert_same(gpu, ref, f"case {case}: bits={bits} c={c} n={n} {w}x{h} abs={absolute} lo={lo} "
                                  f"hi={hi} bg={bg} sel={sel} w={weights} fade={fade} slots={slots} "
                                  f"crop={crop is not None} sample={sample}")
        finally:
            if c >= 3:
                cpa.tile_shapes_set(ctx, bits // 8, c, [])


def test_random_two_pixel_tiles_with_crops(ctx):
    """24 configurations of the two-pixel 16-bit path from a fixed seed: 100-256 frames,
    several tiles per row, per-frame crop origins of up to +-8 pixels (every row alignment of
    the pair loads), 3 and 4 channels, absolute thresholds in the modes the default workload
    uses and beyond."""
    rng = np.random.default_rng(20261208)
    for case in range(24):
        c = int(rng.choice([3, 4]))
        n = int(rng.integers(100, 257))
        h, w = int(rng.integers(18, 26)), int(rng.integers(60, 300))
        x = make_stack(int(rng.integers(1 << 30)), n, h, w, c, bits=16, outlier_frac=float(rng.uniform(0, 0.2)))
        offs = rng.integers(-8, 9, (n, 2)).tolist()
        offs[0] = [0, 0]
        crop = O.crop_create(offs, w, h)
        bg, sel = int(rng.choice([0, 0, 1, 2, 3])), int(rng.choice([4, 4, 0, 1, 2, 3, 5]))
        lo = float(rng.uniform(0.02, 0.1))
        gpu, ref = run_both(ctx, x, True, lo, lo + float(rng.uniform(0, 0.2)), bg, sel, crop=crop,
                            seed=int(rng.integers(1 << 31)))
        assert_same(gpu, ref, f"case {case}: c={c} n={n} {w}x{h} bg={bg} sel={sel}")


def test_random_simple_configurations(ctx):
    """120 darker/lighter configurations from a fixed seed: sample width, 1-4 channels,
    1-200 frames, width, weights, fade, frame selection and crop, against the oracle."""
    rng = np.random.default_rng(20261204)
    for case in range(120):
        bits = int(rng.choice([8, 16]))
        c = int(rng.integers(1, 5))
        n = int(rng.integers(1, 200))
        w, h = int(rng.integers(2, 70)), int(rng.integers(1, 7))  # make_stack writes two tie pixels
        x = make_stack(int(rng.integers(1 << 30)), n, h, w, c, bits=bits, outlier_frac=float(rng.uniform(0, 0.5)))
        st = cpa.Stack.from_array(ctx, x)
        darker = bool(rng.random() < 0.5)
        weights = tuple(float(v) for v in rng.choice([1.0, 0.5, 2.0, 0.0, -1.0], 4))
        slots = None
        if n > 3 and rng.random() < 0.3:
            slots = sorted(rng.choice(n, int(rng.integers(2, n)), replace=False).tolist())
        m = n if slots is None else len(slots)
        fade, cfade, ofade = None, None, None
        if rng.random() < 0.4:
            keys = sorted(rng.choice(max(n, 4), int(rng.integers(2, 5)), replace=False).tolist())
            fade = (bool(rng.random() < 0.5), bool(rng.random() < 0.5),
                    [(int(k), float(rng.choice([0.0, 0.25, 0.5, 1.0, float(rng.random())]))) for k in keys])
            cfade = cpa.Fade(*fade)
            ofade = O.Fade(fade[0], fade[1], [f for f, _ in fade[2]], [v for _, v in fade[2]])
        crop = None
        if w > 5 and h > 5 and rng.random() < 0.3:
            offs = rng.integers(-2, 3, (m, 2)).tolist()
            offs[0] = [0, 0]
            cr = O.crop_create(offs, w, h)
            if cr is not None and cr[1] > 0 and cr[2] > 0:
                crop = cr
        img = cpa.SimpleProcessor(weights, cfade, darker).process(st, image_indices=slots, crop=crop)
        ref = O.simple(x, weights, ofade, darker, slots=slots, crop_xy=None if crop is None else crop[0],
                       out_w=None if crop is None else crop[1], out_h=None if crop is None else crop[2])
        assert np.array_equal(img, ref), f"case {case}: bits={bits} c={c} n={n} {w}x{h} darker={darker} " \
                                         f"w={weights} fade={fade} slots={slots} crop={crop is not None}"
        st.close()


def test_tile_shapes_set_rejects_non_candidates(ctx):
    with pytest.raises(cpa.ChronoError) as e:
        cpa.tile_shapes_set(ctx, 1, 3, [None, (8, 64, 1), None, None, None, None])
    assert e.value.code == cpa.CP_ERR_INVALID_ARG


@pytest.mark.parametrize("bits,n", [(8, 513), (8, 700), (8, 1300), (16, 300), (16, 520)])
def test_outlier_tall_stacks(ctx, bits, n):
    """Tall stacks: 8-bit up to 1280 frames on the packed-byte kernel (16 lanes per pixel),
    8-bit beyond it, 16-bit over 512 frames and 16-bit average modes whose integer sums could
    leave f32's exact range (N * 65535 >= 2^24) on the generic kernel; every mode family stays
    exact."""
    x = make_stack(2100 + n + bits, n, 4, 9, 3, bits=bits, outlier_frac=0.1)
    for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3),
                                      (True, 0.05, 0.2, 1, 2), (False, 1.0, 1.5, 3, 5)):
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel)
        assert_same(gpu, ref, f"bits={bits} n={n} abs={absolute} bg={bg} sel={sel}")


@pytest.mark.parametrize("bits,c", [(16, 3), (16, 1), (8, 2), (16, 4)])
def test_outlier_generic_kernel_modes(ctx, bits, c):
    """The generic kernel forced on 16-bit and 1/2/4-channel stacks across the mode families
    (it is also what 16-bit 1- and 2-channel stacks always use)."""
    x = make_stack(2300 + bits + c, 23, 6, 11, c, bits=bits)
    for bg, sel in ((0, 0), (1, 1), (2, 3), (3, 4), (0, 5), (2, 2)):
        for absolute, lo, hi in ((True, 0.05, 0.2), (False, 2.0, 4.0)):
            gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, force_generic=True)
            assert_same(gpu, ref, f"bits={bits} c={c} bg={bg} sel={sel} abs={absolute}")


@pytest.mark.parametrize("bits", [8, 16])
def test_outlier_pinned_blend_scenarios(ctx, bits):
    """The oracle's blend pins run on the GPU too: a zero fade (output = background) under
    every background mode, and per-frame fades of 1/4, 1/2, 3/4 with full blend values in the
    average, forward and backward modes."""
    n = 20
    zero = (False, True, [(0, 0.0), (1, 0.0)])
    rng = np.random.default_rng(37)
    dyadic = (False, True, [(t, float(v)) for t, v in enumerate(rng.choice([0.25, 0.5, 0.75], n))])
    for frac in (0.25, 0.04):
        x = make_stack(int(frac * 100) + bits, n, 9, 13, 3, bits, outlier_frac=frac)
        for bg in (0, 1, 2, 3):
            gpu, ref = run_both(ctx, x, True, 0.05, 0.2, bg, 2, fade=zero)
            assert_same(gpu, ref, f"zero fade bg {bg} frac {frac}")
        for sel in (3, 4, 5):
            gpu, ref = run_both(ctx, x, True, 0.05, 0.05, 0, sel, fade=dyadic)
            assert_same(gpu, ref, f"dyadic fade sel {sel} frac {frac}")


def test_outlier_fade_windows_crop_sample(ctx):
    x = make_stack(900, 30, 12, 16, 3)
    fade = (False, True, [(2, 0.2), (10, 1.0), (20, 0.4)])
    slots = list(range(5, 25))
    gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 0, 4, fade=fade, slots=slots)
    assert_same(gpu, ref, "fade abs window")
    fade_rel = (True, False, [(0, 0.0), (7, 1.0)])
    gpu, ref = run_both(ctx, x, True, 0.05, 0.3, 2, 5, fade=fade_rel, slots=slots)
    assert_same(gpu, ref, "fade rel repeat")
    rng = np.random.default_rng(1)
    crop_xy = rng.integers(0, 4, (30, 2)).astype(np.int32)
    gpu, ref = run_both(ctx, x, False, 2.0, 4.0, 1, 2, crop=(crop_xy, 12, 8))
    assert_same(gpu, ref, "crop")
    sample = [0, 3, 4, 9, 11, 12, 20, 28, 29]
    gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 3, 3, sample=sample)
    assert_same(gpu, ref, "sample")
    gpu, ref = run_both(ctx, x, False, 2.0, 4.0, 1, 4, sample=sample)
    assert_same(gpu, ref, "sample rel")


def test_outlier_rel_too_few_frames(ctx):
    x = make_stack(1, 2, 4, 4, 3)
    st = cpa.Stack.from_array(ctx, x)
    proc = cpa.OutlierProcessor(cpa.Threshold.rel(2.0, 4.0), 0, 2)
    with pytest.raises(cpa.ChronoError) as e:
        proc.process(st)
    assert e.value.code == cpa.CP_ERR_TOO_FEW_FRAMES_REL


# ------------------------------------------------------------------ darker / lighter

@pytest.mark.parametrize("darker", [True, False])
@pytest.mark.parametrize("bits,c", [(8, 3), (8, 4), (16, 3), (8, 1)])
def test_simple_parity(ctx, darker, bits, c):
    x = make_stack(40 + bits + c, 19, 10, 23, c, bits=bits, outlier_frac=0.3)
    st = cpa.Stack.from_array(ctx, x)
    for w, fade in [((1, 1, 1, 1), None), ((0.2, 0.7, 0.1, 1.0), (False, True, [(0, 0.3), (18, 1.0)])),
                    ((1, -1, 0.5, 0), (True, False, [(0, 1.0), (4, 0.0)]))]:
        cfade = cpa.Fade.none() if fade is None else cpa.Fade(*fade)
        ofade = O.Fade.none() if fade is None else O.Fade(fade[0], fade[1], [f for f, _ in fade[2]],
                                                            [v for _, v in fade[2]])
        img = cpa.SimpleProcessor(w, cfade, darker).process(st)
        ref = O.simple(x, w, ofade, darker)
        assert np.array_equal(img, ref), (darker, bits, c, w)
        slots = list(range(3, 15))
        img = cpa.SimpleProcessor(w, cfade, darker).process(st, image_indices=slots)
        ref = O.simple(x, w, ofade, darker, slots=slots)
        assert np.array_equal(img, ref)


@pytest.mark.parametrize("bits,c", [(8, 1), (8, 2), (8, 3), (8, 4), (16, 1), (16, 2), (16, 3), (16, 4)])
def test_simple_widths_and_crops(ctx, bits, c):
    """Every row alignment the lanes' dword reads can meet: widths 1..9 and 61, crop origins
    that move each frame's row start by one sample, and a partial last lane."""
    rng = np.random.default_rng(60 + bits + c)
    n, h = 11, 5
    for w in list(range(1, 10)) + [61]:
        x = make_stack(int(rng.integers(1 << 30)), n, h, w + 3, c, bits=bits, outlier_frac=0.3)
        st = cpa.Stack.from_array(ctx, x)
        offs = [(int(rng.integers(0, 4)), int(rng.integers(0, 2))) for _ in range(n)]
        offs[0] = (0, 0)
        crop = cpa.Crop.create([(dx - 3, dy - 1) for dx, dy in offs], w + 3, h)
        ocrop = O.crop_create([(dx - 3, dy - 1) for dx, dy in offs], w + 3, h)
        for darker in (True, False):
            img = cpa.SimpleProcessor((1, 0.5, 2, 1)[:4], None, darker).process(st)
            assert np.array_equal(img, O.simple(x, (1, 0.5, 2, 1), None, darker)), (w, darker)
            if crop is not None:
                img = cpa.SimpleProcessor(darker=darker).process(st, crop=crop)
                ref = O.simple(x, darker=darker, crop_xy=ocrop[0], out_w=ocrop[1], out_h=ocrop[2])
                assert np.array_equal(img, ref), ("crop", w, darker, offs)
        st.close()


# --------------------------------------------------------------------------- shake

def textured(n, h, w, shifts, c=3, bits=8, seed=0):
    rng = np.random.default_rng(seed)
    hi = (1 << bits) - 1
    world = rng.integers(0, hi + 1, (h + 40, w + 40, c))
    out = np.zeros((n, h, w, c), np.int64)
    for t, (dx, dy) in enumerate(shifts):
        out[t] = world[20 - dy:20 - dy + h, 20 - dx:20 - dx + w]
    return out.astype(np.uint8 if bits == 8 else np.uint16)


@pytest.mark.parametrize("bits", [8, 16])
def test_shake_parity(ctx, bits):
    shifts = [(0, 0), (2, -1), (-3, 3), (1, 0), (0, -2), (4, 4), (-4, -1)]
    x = textured(len(shifts), 48, 64, shifts, bits=bits)
    st = cpa.Stack.from_array(ctx, x)
    anchors = [(20, 16), (40, 30)]
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 5, 4)
    ref = O.shake(x, anchors, 5, 4)
    assert np.array_equal(off, ref)
    assert [tuple(o) for o in off] == shifts


def test_shake_u8_wrapping_sum_matches(ctx):
    # bright/dark contrast with big windows drives the i32 SSD past 2^31 (wraps like the
    # reference's release build)
    shifts = [(0, 0), (1, 1), (-2, 0)]
    x = textured(3, 120, 120, shifts, bits=8, seed=3)
    x[1:] = 255 - x[1:]
    st = cpa.Stack.from_array(ctx, x)
    anchors = [(60, 60), (50, 55), (70, 62)]
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 30, 6)
    assert np.array_equal(off, O.shake(x, anchors, 30, 6))


@pytest.mark.parametrize("c,bits,r,s", [(1, 8, 4, 1), (2, 16, 3, 2), (4, 8, 6, 3), (4, 16, 9, 7),
                                         (3, 16, 12, 5), (1, 16, 20, 0), (3, 8, 0, 4), (3, 16, 10, 32),
                                         (3, 8, 3, 40), (4, 16, 60, 12)])
def test_shake_unstructured_frames(ctx, c, bits, r, s):
    """Random frames: every candidate's sum differs, so the arg-min exposes any wrong term.
    Covers partial offset strips (2s+1 not a multiple of 5), several window-row chunks,
    candidate rows split over workgroups (large s, or large r with s) and 1-4 channels."""
    rng = np.random.default_rng(100 + c * 7 + r + s)
    hi = (1 << bits) - 1
    h, w = 2 * (r + s) + 17, 2 * (r + s) + 23
    x = rng.integers(0, hi + 1, (6, h, w, c)).astype(np.uint8 if bits == 8 else np.uint16)
    anchors = [(r + s, r + s), (w - 1 - r - s, h - 1 - r - s), (w // 2, h // 2)]
    st = cpa.Stack.from_array(ctx, x)
    off = cpa.ShakeAnalyzer().analyze(st, anchors, r, s)
    assert np.array_equal(off, O.shake(x, anchors, r, s))


def test_shake_large_anchor_region(ctx):
    """r=50, s=10 on RGB16: a 121x121 search region, split over several workgroups' LDS."""
    shifts = [(0, 0), (3, -2), (-7, 5), (10, 10), (-10, -9)]
    x = textured(len(shifts), 150, 170, shifts, bits=16, seed=11)
    anchors = [(70, 70), (100, 75)]
    st = cpa.Stack.from_array(ctx, x)
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 50, 10)
    assert np.array_equal(off, O.shake(x, anchors, 50, 10))
    assert [tuple(o) for o in off] == shifts


@pytest.mark.parametrize("bits", [8, 16])
def test_shake_ties_keep_first_offset(ctx, bits):
    """Constant frames: every candidate sums to 0 and min_by_key keeps offset (-s, -s)."""
    x = np.full((4, 40, 40, 3), 200, np.uint8 if bits == 8 else np.uint16)
    st = cpa.Stack.from_array(ctx, x)
    off = cpa.ShakeAnalyzer().analyze(st, [(20, 20)], 6, 7)
    assert np.array_equal(off, O.shake(x, [(20, 20)], 6, 7))
    assert off[1:].tolist() == [[-7, -7]] * 3


def test_shake_frame_selection_single_frame_and_no_anchors(ctx):
    """Frame slots in any order (the reference's image list after --frames), one frame only
    (offset (0, 0)), and no anchors (every sum 0: min_by_key keeps offset (-s, -s))."""
    shifts = [(0, 0), (2, -1), (-3, 3), (1, 0), (0, -2), (4, 4)]
    x = textured(len(shifts), 48, 64, shifts, bits=16, seed=4)
    st = cpa.Stack.from_array(ctx, x)
    anchors = [(20, 16), (40, 30)]
    slots = [3, 0, 5, 1]
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 5, 4, frames=slots)
    assert np.array_equal(off, O.shake(x, anchors, 5, 4, slots=slots))
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 5, 4, frames=[2])
    assert off.tolist() == [[0, 0]] and np.array_equal(off, O.shake(x, anchors, 5, 4, slots=[2]))
    off = cpa.ShakeAnalyzer().analyze(st, [], 5, 4)
    assert np.array_equal(off, O.shake(x, [], 5, 4))
    assert off[1:].tolist() == [[-4, -4]] * (len(shifts) - 1)


def test_shake_random_configurations(ctx):
    """60 shake searches from a fixed seed: sample width, 1-4 channels, 2-9 frames (some
    through a frame selection), 1-4 anchors anywhere they fit, r in 0..14, s in 0..11."""
    rng = np.random.default_rng(20261205)
    for case in range(60):
        bits = int(rng.choice([8, 16]))
        c = int(rng.integers(1, 5))
        r, s = int(rng.integers(0, 15)), int(rng.integers(0, 12))
        n = int(rng.integers(2, 10))
        h, w = 2 * (r + s) + int(rng.integers(1, 20)), 2 * (r + s) + int(rng.integers(1, 30))
        hi = (1 << bits) - 1
        x = rng.integers(0, hi + 1, (n, h, w, c)).astype(np.uint8 if bits == 8 else np.uint16)
        anchors = [(int(rng.integers(r + s, w - r - s)), int(rng.integers(r + s, h - r - s)))
                   for _ in range(int(rng.integers(1, 5)))]
        slots = None
        if rng.random() < 0.3:
            slots = rng.permutation(n)[: int(rng.integers(1, n + 1))].tolist()
        st = cpa.Stack.from_array(ctx, x)
        off = cpa.ShakeAnalyzer().analyze(st, anchors, r, s, frames=slots)
        ref = O.shake(x, anchors, r, s, slots=slots)
        assert np.array_equal(off, ref), f"case {case}: bits={bits} c={c} n={n} {w}x{h} r={r} s={s} " \
                                         f"anchors={anchors} slots={slots}"
        st.close()


def test_shake_anchor_errors(ctx):
    x = textured(2, 30, 30, [(0, 0), (1, 1)])
    st = cpa.Stack.from_array(ctx, x)
    with pytest.raises(cpa.ChronoError) as e:
        cpa.ShakeAnalyzer().analyze(st, [(4, 4)], 3, 3)
    assert e.value.code == cpa.CP_ERR_ANCHOR_OUT_OF_RANGE


def test_shake_recovers_synthetic_translations(ctx):
    """Static textured scene: the measured offsets are the seeded per-frame translations."""
    scene = cpa.Scene(seed=5, width=96, height=72, bytes_per_sample=2, n_frames=12, base_lo=8000,
                      base_hi=50000, sigma=200, n_disks=0, shake_max=3, texture_amp=3000)
    x = scene.generate()
    st = cpa.Stack.from_array(ctx, x)
    off = cpa.ShakeAnalyzer().analyze(st, [(24, 20), (70, 50)], 6, 4)
    assert np.array_equal(off, scene.shake_truth())
    assert np.array_equal(off, O.shake(x, [(24, 20), (70, 50)], 6, 4))


def test_shake_crop_outlier_pipeline(ctx):
    """Shake offsets -> Crop::create -> outlier composite on the cropped stack (moving disks
    may cross the anchors, so offsets are compared with the oracle, not the seed)."""
    scene = cpa.Scene(seed=5, width=96, height=72, bytes_per_sample=2, n_frames=12, base_lo=8000,
                      base_hi=50000, sigma=200, n_disks=2, disk_r_min=6, disk_r_max=10,
                      disk_val_lo=2000, disk_val_hi=3000, shake_max=3, texture_amp=3000)
    x = scene.generate()
    st = cpa.Stack.from_array(ctx, x)
    anchors = [(24, 20), (70, 50)]
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 6, 4)
    assert np.array_equal(off, O.shake(x, anchors, 6, 4))
    crop = cpa.Crop.create(off, 96, 72)
    ocrop = O.crop_create(off, 96, 72)
    assert crop is not None and np.array_equal(crop[0], ocrop[0]) and crop[1:] == ocrop[1:]
    gpu, ref = run_both(ctx, x, True, 0.05, 0.2, 0, 4, crop=crop)
    assert_same(gpu, ref, "shake crop")


# ------------------------------------------------------------------ synthetic scenes

def test_device_generator_matches_host(ctx):
    for bps in (1, 2):
        scene = cpa.Scene(seed=9, width=200, height=150, bytes_per_sample=bps, n_frames=5,
                          base_lo=40 if bps == 1 else 8000, base_hi=200 if bps == 1 else 50000,
                          sigma=3.0 if bps == 1 else 200.0, n_disks=2, disk_r_min=8, disk_r_max=20,
                          disk_val_lo=15, disk_val_hi=25, blob_permille=300, shake_max=5, texture_amp=30)
        host = scene.generate()
        st = cpa.Stack(ctx, 200, 150, 3, bps, capacity=5)
        scene.generate_device(st.device_ptr, 0, 5, stream=ctx.stream)
        ctx.synchronize()
        st.commit(5)
        assert np.array_equal(st.download(), host), bps


def test_shake_bench_radii_mid_size(ctx):
    """Anchor radius 32, search radius 12, five anchors, as in the 4K benchmark."""
    scene = cpa.Scene(seed=2, width=640, height=480, bytes_per_sample=2, n_frames=6, base_lo=8000,
                      base_hi=50000, sigma=200, n_disks=0, shake_max=8, texture_amp=3000)
    x = scene.generate()
    st = cpa.Stack.from_array(ctx, x)
    anchors = [(60, 60), (580, 60), (320, 240), (60, 420), (580, 420)]
    off = cpa.ShakeAnalyzer().analyze(st, anchors, 32, 12)
    ref = O.shake(x, anchors, 32, 12)
    assert np.array_equal(ref, scene.shake_truth())
    assert np.array_equal(off, ref), (off.tolist(), ref.tolist())
