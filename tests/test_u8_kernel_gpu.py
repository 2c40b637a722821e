"""The 8-bit packed-byte outlier kernel against the oracle: tall stacks up to 1280 frames
(16 lanes per pixel), 1- and 2-channel stacks, `--sample` subsets ranked in their own block,
every tile shape, partial tiles, weights, crops and frame selections.  Bit-exact."""
import numpy as np
import pytest

import chrono_photo_amd as cpa
import oracle_py as O
from test_parity_gpu import MODES, assert_same, make_stack, run_both

pytestmark = pytest.mark.gpu

# packed-byte choices of every bucket (lanes per pixel group, frames per lane, pixels per lane,
# kernel): the two sets differ in the 512-frame bucket (four pixels or one pixel per lane)
U8_CHOICES = [[(1, 16, 1, 1), (1, 32, 1, 1), (2, 32, 1, 1), (4, 32, 1, 1), (8, 32, 1, 1), (16, 32, 4, 1),
               (32, 24, 4, 1), (32, 32, 4, 1), (32, 40, 4, 1)],
              [(1, 16, 1, 1), (1, 32, 1, 1), (2, 32, 1, 1), (4, 32, 1, 1), (8, 32, 1, 1), (16, 32, 1, 1),
               (32, 24, 4, 1), (32, 32, 4, 1), (32, 40, 4, 1)]]


def _tall(seed, n, h, w, c, frac=0.1):
    x = make_stack(seed, n, h, w, c, bits=8, outlier_frac=frac)
    x[:, 1, 2:6, :] = 200                               # constant pixels
    x[: (n - 1) // 2, 2, 2:6, :] = 10                   # outliers over the first half
    x[:, 3, 2:6, :] = 255 - (np.arange(n) % 3)[:, None, None]  # at the top of the range
    return x


@pytest.mark.parametrize("n", [600, 777, 1024, 1280])
@pytest.mark.parametrize("c", [3, 4, 1])
def test_tall_stacks_all_modes(ctx, n, c):
    """Every background x selection mode, absolute and relative thresholds, over stacks of
    600-1280 frames (partial last lane blocks at 600 and 777)."""
    x = _tall(3100 + n + c, n, 4, 19, c)
    for bg, sel in MODES:
        absolute = (bg + sel) % 2 == 0
        lo, hi = (0.05, 0.2) if absolute else (2.0, 4.0)
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel)
        assert_same(gpu, ref, f"n={n} c={c} bg={bg} sel={sel} abs={absolute}")


@pytest.mark.parametrize("c", [1, 2])
@pytest.mark.parametrize("n", [1, 3, 5, 16, 33, 64, 200, 512])
def test_one_and_two_channel_stacks(ctx, c, n):
    x = make_stack(3300 + n + c, n, 5, 23, c, bits=8)
    for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3), (True, 0.1, 0.1, 1, 2),
                                      (False, 1.0, 1.5, 3, 5), (True, 0.02, 0.3, 2, 0), (False, 0.5, 2.0, 0, 1)):
        if not absolute and n < 3:
            continue
        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, weights=(1.0, 0.5, 1, 1))
        assert_same(gpu, ref, f"c={c} n={n} abs={absolute} bg={bg} sel={sel}")


@pytest.mark.parametrize("n,s", [(40, 3), (40, 16), (300, 64), (300, 128), (1200, 64), (1280, 128), (1200, 1),
                                 (300, 200)])
@pytest.mark.parametrize("c", [3, 1, 4])
def test_sample_subsets(ctx, n, s, c):
    """`--sample`: medians and quartiles over a sorted subset of the positions, distances over
    all of them; subsets of up to 128 frames take their own block of the packed-byte kernel,
    larger ones the generic kernel."""
    rng = np.random.default_rng(n * 7 + s + c)
    x = _tall(3500 + n + s + c, n, 4, 21, c, frac=0.15)
    sample = sorted(rng.choice(n, s, replace=False).tolist())
    for bg, sel in ((0, 4), (2, 3), (1, 2), (3, 5), (0, 0), (2, 1)):
        for absolute, lo, hi in ((True, 0.05, 0.2), (False, 2.0, 4.0)):
            if not absolute and s < 3:
                continue
            gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, sample=sample)
            assert_same(gpu, ref, f"n={n} s={s} c={c} bg={bg} sel={sel} abs={absolute}")


@pytest.mark.parametrize("c", [3, 4])
def test_every_packed_byte_shape(ctx, c):
    """The packed-byte choice pinned in every bucket, at frame counts that fill each shape and
    one frame short of it."""
    try:
        for pick, choice in enumerate(U8_CHOICES):
            cpa.tile_shapes_set(ctx, 1, c, choice)
            for nmax in cpa.TILE_BUCKETS[(5 if pick else 0):(6 if pick else None)]:
                for n in (nmax, nmax - 1):
                    x = _tall(3700 + n + c, n, 4, 70, c, frac=0.05)
                    for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3)):
                        gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel)
                        assert_same(gpu, ref, f"c={c} n={n} pick={pick} abs={absolute}")
    finally:
        cpa.tile_shapes_set(ctx, 1, c, [])


def test_tall_weights_crops_and_selections(ctx):
    """Zero, negative and fractional weights, per-frame crops and a frame selection over a
    tall stack, with fades on the selected frames."""
    n, h, w = 900, 8, 41
    x = _tall(3900, n, h, w, 3, frac=0.08)
    rng = np.random.default_rng(39)
    slots = sorted(rng.choice(n, 700, replace=False).tolist())
    offs = rng.integers(-2, 3, (len(slots), 2)).tolist()
    offs[0] = [0, 0]
    crop = O.crop_create(offs, w, h)
    fade = (False, True, [(slots[0], 0.25), (slots[0] + 300, 1.0), (slots[0] + 600, 0.5)])
    for weights in ((1, 1, 1, 1), (0.5, 2.0, 0.0, 1.0), (1.0, -1.0, 1.0, 1.0)):
        for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3), (True, 0.03, 0.1, 1, 5)):
            gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, weights=weights, fade=fade, slots=slots,
                                crop=crop)
            assert_same(gpu, ref, f"weights={weights} abs={absolute} bg={bg} sel={sel}")


@pytest.mark.parametrize("c,shift", [(3, 0), (3, 1), (4, 0), (1, 3), (2, 2)])
def test_borrowed_stack_ending_at_its_last_frame(ctx, c, shift):
    """A borrowed allocation ends at its last frame and may start anywhere: composites whose
    wide loads would read past that end (or 4-channel stacks off a dword boundary) take the
    generic kernel, the rest the packed-byte kernel; both exact."""
    from chrono_photo_amd import hip as H
    for n in (40, 700):
        x = np.ascontiguousarray(_tall(4100 + n + c, n, 5, 27, c))
        buf = H.DeviceBuffer(x.nbytes + shift)
        H.check(H.hip().hipMemcpy(buf.ptr + shift, x.ctypes.data, x.nbytes, H.hipMemcpyHostToDevice), "hipMemcpy")
        st = cpa.Stack(ctx, 27, 5, c, 1, device_ptr=buf.ptr + shift, n_frames=n)
        for sel in (2, 4):
            proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), 0, sel)
            img, mask = proc.process(st)
            oimg, omask, _ = O.outlier(x, O.threshold(True, 0.05, 0.2, 1), 0, sel)
            assert np.array_equal(img, oimg), (c, shift, n, sel)
            assert np.array_equal(mask, omask), (c, shift, n, sel)
        st.close()
        buf.free()


@pytest.mark.parametrize("n", [300, 1100])
def test_four_pixel_tiles_with_crops(ctx, n):
    """Per-frame crop origins of up to +-3 pixels move every frame's 4-pixel group to every
    byte alignment of its 16-byte load; rows with partial last tiles."""
    rng = np.random.default_rng(n)
    h, w = 7, 75
    x = _tall(4300 + n, n, h, w, 3, frac=0.1)
    offs = rng.integers(-3, 4, (n, 2)).tolist()
    offs[0] = [0, 0]
    crop = O.crop_create(offs, w, h)
    try:
        cpa.tile_shapes_set(ctx, 1, 3, U8_CHOICES[0])
        for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3), (True, 0.03, 0.1, 1, 2)):
            gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, crop=crop, seed=5)
            assert_same(gpu, ref, f"n={n} abs={absolute} bg={bg} sel={sel}")
    finally:
        cpa.tile_shapes_set(ctx, 1, 3, [])


SORT_CHOICE = [(1, 16, 1, 2), (1, 32, 1, 2), (1, 64, 1, 2)]


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_sorting_network_ranks(ctx, c):
    """The one-lane-per-pixel shapes with sorting-network ranks (buckets of up to 16, 32 and
    64 frames) under every mode family, absolute and relative, with ties, constant pixels and
    pads (frame counts below the network's size)."""
    try:
        cpa.tile_shapes_set(ctx, 1, c, SORT_CHOICE)
        for n in (1, 2, 3, 5, 16, 17, 31, 32, 33, 50, 64):
            x = make_stack(4500 + n + c, n, 4, 29, c, bits=8, outlier_frac=0.2)
            for bg, sel in MODES[::5] + [(2, 3), (3, 5)]:
                for absolute, lo, hi in ((True, 0.05, 0.2), (False, 2.0, 4.0)):
                    if not absolute and n < 3:
                        continue
                    gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel, weights=(1.0, 2.0, 0.5, 1.0))
                    assert_same(gpu, ref, f"c={c} n={n} bg={bg} sel={sel} abs={absolute}")
    finally:
        cpa.tile_shapes_set(ctx, 1, c, [])


# relative-threshold choices per bucket from 64 frames up: the 16-bit-halves byte search and
# the packed-byte byte search, both of which search the quartiles next to the median
QUARTILE_CHOICES = [[None, None, (2, 32, 1), (4, 32, 1), (8, 32, 1), (8, 64, 1)],
                    [None, None, (2, 32, 1, 1), (4, 32, 1, 1), (8, 32, 1, 1), (16, 32, 1, 1)]]


@pytest.mark.parametrize("c", [3, 1])
def test_quartile_windows(ctx, c):
    """Relative thresholds over pixels whose quartiles lie within 15 levels of the median (the
    windowed 4-step searches), pixels near 0 and 255 (windows clipped at the range ends),
    pixels with quartiles far from the median in some lanes of a wave (every wave falls back
    to the full searches), and rows of each kind alone, for 64-512 frames."""
    rng = np.random.default_rng(4100 + c)
    for choice in QUARTILE_CHOICES if c == 3 else QUARTILE_CHOICES[1:]:
        try:
            cpa.tile_shapes_set(ctx, 1, c, [], choice)
            for n in (64, 65, 128, 200, 256, 300, 512):
                h, w = 6, 70
                centre = np.array([128, 5, 250, 128, 14, 241])[:, None, None]  # one row per kind
                x = centre[None] + np.rint(rng.normal(0, 3, (n, h, w, c))).astype(int)
                wide = rng.integers(0, 256, (n, w, c))
                x[:, 3] = wide                                  # quartiles far from the median
                x[:, 5, : w // 2] = wide[:, : w // 2]           # half a row: waves that mix both
                hit = rng.random((n, h, w)) < 0.05
                x = np.where(hit[..., None], rng.integers(0, 256, x.shape), x)
                x = np.clip(x, 0, 255).astype(np.uint8)
                for bg, sel in ((0, 4), (2, 3), (1, 0)):
                    gpu, ref = run_both(ctx, x, False, 2.0, 4.0, bg, sel)
                    assert_same(gpu, ref, f"c={c} n={n} bg={bg} sel={sel} choice={choice}")
        finally:
            cpa.tile_shapes_set(ctx, 1, c, [])


@pytest.mark.parametrize("c", [3, 1])
def test_median_windows(ctx, c):
    """8-bit medians searched in a 16-value window around an estimate, under every candidate
    tile shape of every bucket: in each group of 8 frames the first 4 sit D above the rest
    (two of each triple of frames the estimate reads), so the estimate lands about D above a
    median that stays with the lower half.  D from 0 to 29 per column puts the median inside
    the window or past its edge (the 8-step search); rows near 0 and 255 clip the window at
    the range ends; stacks below 12 frames per lane block take the 8-step search directly."""
    from test_parity_gpu import CANDIDATES, PICKS
    rng = np.random.default_rng(4300 + c)
    h, w = 3, 60
    if c == 3:
        sets = [[cands[min(pick, len(cands) - 1)] for cands in CANDIDATES[8]] for pick in PICKS]
    else:
        sets = U8_CHOICES + [SORT_CHOICE, [None, None, (1, 64, 1, 1)]]
    for shapes in sets:
        try:
            cpa.tile_shapes_set(ctx, 1, c, shapes)
            for n in (8, 16, 33, 64, 100, 256, 300, 512, 640, 1280):
                base = np.array([120, 2, 248])[:, None, None]
                d = (np.arange(w) % 30)[None, :, None]
                x = base[None] + rng.integers(-1, 2, (n, h, w, c))
                shifted = (np.arange(n) % 8) < 4
                x = np.where(shifted[:, None, None, None], x + d[None], x)
                x = np.clip(x, 0, 255).astype(np.uint8)
                for absolute, lo, hi, bg, sel in ((True, 0.05, 0.2, 0, 4), (False, 2.0, 4.0, 2, 3)):
                    gpu, ref = run_both(ctx, x, absolute, lo, hi, bg, sel)
                    assert_same(gpu, ref, f"c={c} n={n} shapes={shapes} abs={absolute}")
        finally:
            cpa.tile_shapes_set(ctx, 1, c, [])


@pytest.mark.parametrize("n", [40, 600, 1201, 1279])
def test_bound_check_partial_lanes(ctx, n):
    """Lanes whose frame block ends in pads: the bound check that skips outlier-free lanes
    takes its maxima over the real frames only.  Quiet pixels with a single outlier in the
    last real frame (often the only real frame of its lane), the first frame or a middle one,
    under absolute thresholds (every selection mode) on the tall and the short tiles."""
    rng = np.random.default_rng(4500 + n)
    h, w, c = 2, 40, 3
    x = 100 + rng.integers(-1, 2, (n, h, w, c))
    x[n - 1, :, 0:8] = 250
    x[0, :, 8:16] = 5
    x[n // 2, :, 16:24] = 240
    x[n - 1, 1, 24:32, 1] = 0                            # one channel only
    x = x.astype(np.uint8)
    for bg, sel in ((0, 0), (0, 1), (0, 2), (1, 4), (2, 3), (3, 5)):
        gpu, ref = run_both(ctx, x, True, 0.05, 0.2, bg, sel)
        assert_same(gpu, ref, f"n={n} bg={bg} sel={sel}")


def test_sorting_tile_stored_distances(ctx):
    """The 64-frame sorting tile keeps up to 16 outlier distances per pixel for the average
    mode's finish and re-reads the samples beyond that: columns with exactly k outliers, k
    from 0 to 30 (half of them far below, half far above a tight majority, so the quartiles
    stay with the majority and every far frame is an outlier), at frames spread over the
    stack, for full and partial (48-frame) stacks.  The majority varies in channel 0 only
    (ten levels, so its quartiles lie 5 apart and no majority frame reaches 2 IQR)."""
    rng = np.random.default_rng(4600)
    ks = [0, 1, 2, 15, 16, 17, 18, 30]
    for n in (64, 48):
        h, w, c = 2, 8 * len(ks), 3
        x = np.full((n, h, w, c), 100)
        x[..., 0] = 96 + rng.integers(0, 10, (n, h, w))
        for i, kk in enumerate(ks):
            kk = min(kk, n // 2 - 2)
            for col in range(8 * i, 8 * i + 8):
                frames = rng.choice(n, kk, replace=False)
                x[frames[: kk // 2], :, col] = 5 + rng.integers(0, 3, (kk // 2, h, c))
                x[frames[kk // 2:], :, col] = 245 + rng.integers(0, 3, (kk - kk // 2, h, c))
        x = x.astype(np.uint8)
        try:
            cpa.tile_shapes_set(ctx, 1, c, [], [None, None, (1, 64, 1, 2)])
            for bg in (0, 2, 3):
                gpu, ref = run_both(ctx, x, False, 2.0, 4.0, bg, 3)
                assert_same(gpu, ref, f"n={n} bg={bg}")
                counts = gpu[2]["num_outliers"].reshape(h, w)
                assert {int(v) for v in counts[0, ::8]} >= {0, 1, 15, 16, 17} if n == 64 else True
        finally:
            cpa.tile_shapes_set(ctx, 1, c, [])
