"""Pins of the CPU oracle (oracle/chrono_oracle.c) against the reference's own known
answers, closed forms, invariants and brute force.  CPU only."""
import numpy as np
import pytest

import oracle_py as O

F32 = np.float32


def rround(x):
    """Rust f32::round (half away from zero), elementwise on float32."""
    x = np.asarray(x, np.float32)
    return np.where(x >= 0, np.floor(x + F32(0.5)), np.ceil(x - F32(0.5))).astype(np.float32)


# --------------------------------------------------------------- order statistics

def test_quartiles_reference_kat():
    # src/chrono.rs:594-605 quartiles_test: [0..6] -> (1.0, 3.0, 5.0)
    d = np.arange(7, dtype=np.uint32)
    assert (O.quantile(d, 0.25), O.median(d), O.quantile(d, 0.75)) == (1.0, 3.0, 5.0)


@pytest.mark.parametrize("bits", [8, 16])
def test_quantiles_match_hyndman_fan_type6(bits):
    rng = np.random.default_rng(7 + bits)
    for n in list(range(3, 80)) + [127, 128, 255, 256, 399]:
        x = np.sort(rng.integers(0, 1 << bits, n)).astype(np.uint32)
        q = np.quantile(x.astype(np.float64), [0.25, 0.5, 0.75], method="weibull")
        got = (O.quantile(x, 0.25), O.median(x), O.quantile(x, 0.75))
        assert got == tuple(float(v) for v in q), (n, got, q)


def test_median_small_and_even():
    assert O.median(np.array([5], np.uint32)) == 5.0
    assert O.median(np.array([1, 4], np.uint32)) == 2.5
    assert O.median(np.array([0, 255, 255], np.uint32)) == 255.0


# ---------------------------------------------------------------- option values

def test_threshold_new_closed_form():
    t = O.threshold(True, 0.05, 0.2)
    assert t.min == F32(F32(0.05) * F32(255.0))
    assert t.max == F32(F32(0.2) * F32(255.0))
    assert t.scale == F32(F32(1.0) / F32((F32(0.2) - F32(0.05)) * F32(255.0)))
    r = O.threshold(False, 2.0, 4.0)
    assert (r.min, r.max, r.scale) == (2.0, 4.0, 0.5)
    w = O.threshold(True, 0.05, 0.2, bps=2)
    assert w.min == F32(F32(0.05) * F32(65535.0))
    lib = O.lib()
    assert lib.or_blend_value(t, 12.0) == 0.0
    assert lib.or_blend_value(t, 60.0) == 1.0
    mid = lib.or_blend_value(t, 30.0)
    assert mid == F32((F32(30.0) - t.min) * t.scale)
    single = O.threshold(True, 0.1, 0.1)  # max omitted: max == min, scale = inf, never used
    assert lib.or_blend_value(single, single.min) == 0.0
    assert lib.or_blend_value(single, np.nextafter(single.min, np.float32(1e9))) == 1.0


def test_fade_table_linear_and_modes():
    tab = O.fade_table([0, 10], [0.0, 1.0])
    assert len(tab) == 11
    for i in range(11):
        assert tab[i] == F32(F32(0.0) + F32(F32(1.0) - F32(0.0)) * F32(i) / F32(10))
    tab = O.fade_table([2, 4, 8], [1.0, 0.0, 0.5])
    assert list(tab[:3]) == [1.0, 0.5, 0.0] and tab[-1] == 0.5 and len(tab) == 7
    clamp = O.Fade(False, True, [2, 4], [1.0, 0.0])
    assert clamp.get(-5) == 1.0 and clamp.get(100) == 0.0 and clamp.get(3) == 0.5
    rep = O.Fade(True, True, [0, 3], [0.0, 0.75])  # table [0, .25, .5, .75]
    assert [rep.get(f) for f in (-1, 4, 5, 8)] == [0.75, 0.0, 0.25, 0.0]
    assert O.Fade.none().get(12345) == 1.0


def test_blend_into_u8_reference_kat():
    # commented reference test src/color.rs:51-64: blend 0 -> 255 by 0.5 gives 128
    assert list(O.blend_into_u8([0] * 4, [255] * 4, 0.5)) == [128] * 4
    assert list(O.blend_into_u8([0] * 4, [255] * 4, 0.0)) == [0] * 4
    assert list(O.blend_into_u8([0] * 4, [255] * 4, 1.0)) == [255] * 4
    assert list(O.blend_into_u8([10], [20], 0.25)) == [13]  # 12.5 rounds away from zero


# ------------------------------------------------------------------ outlier core

def _stack(rng, n, h, w, c, bits=8, outlier_frac=0.15):
    hi = (1 << bits) - 1
    base = rng.integers(hi // 4, 3 * hi // 4, (1, h, w, c))
    noise = rng.integers(-hi // 40 - 1, hi // 40 + 2, (n, h, w, c))
    x = np.clip(base + noise, 0, hi)
    hit = rng.random((n, h, w)) < outlier_frac
    x = np.where(hit[..., None], rng.integers(0, hi + 1, (n, h, w, c)), x)
    return x.astype(np.uint8 if bits == 8 else np.uint16)


def _numpy_distances(x, med, iqr_inv, w, absolute):
    """Vectorised float32 re-evaluation of src/chrono.rs:257-278 in channel order."""
    n, h, wd, c = x.shape
    d = np.zeros((n, h, wd), np.float32)
    for i in range(c):
        wi = F32(w[i])
        if wi == 0:
            continue
        diff = (med[None, :, :, i] - x[..., i].astype(np.float32)).astype(np.float32)
        if absolute:
            a = (wi * diff).astype(np.float32)
        else:
            a = ((wi * iqr_inv[None, :, :, i]).astype(np.float32) * diff).astype(np.float32)
        term = (F32(np.copysign(1.0, wi)) * (a * a).astype(np.float32)).astype(np.float32)
        term = np.where(diff == 0, F32(0), term)
        d = (d + term).astype(np.float32)
    return d


@pytest.mark.parametrize("absolute,bits", [(True, 8), (False, 8), (True, 16), (False, 16)])
def test_outlier_statistics_and_mask_invariants(absolute, bits):
    rng = np.random.default_rng(11 + bits + absolute)
    x = _stack(rng, 23, 6, 7, 3, bits)
    thr = O.threshold(absolute, 0.05, 0.2, bps=bits // 8) if absolute else O.threshold(False, 2.0, 4.0)
    w = (1.0, 0.5, 2.0, 1.0)
    img, mask, warn, st = O.outlier(x, thr, 0, 2, weights=w, stats=True)
    xs = np.sort(x.astype(np.float64), axis=0)
    q1, med, q3 = np.quantile(xs, [0.25, 0.5, 0.75], axis=0, method="weibull")
    assert np.array_equal(st["median"][..., :3], med.astype(np.float32))
    if absolute:
        assert np.all(st["iqr_inv"] == 0)
        iqr_inv = np.zeros_like(med, np.float32)
    else:
        iqr = (q3 - q1).astype(np.float32)
        iqr = np.where(iqr == 0, F32(1), iqr)
        iqr_inv = (F32(1) / iqr).astype(np.float32)
        assert np.array_equal(st["iqr_inv"][..., :3], iqr_inv)
    d = _numpy_distances(x, med.astype(np.float32), iqr_inv, w, absolute)
    out = d >= F32(thr.min) * F32(thr.min)
    assert np.array_equal(st["num_outliers"], out.sum(axis=0))
    # extreme = first maximum among outliers (strict > keeps the earliest frame)
    dm = np.where(out, d, F32(0))
    has = out.any(axis=0)
    first_max = np.argmax(dm, axis=0)
    assert np.array_equal(np.where(has, st["max_index"], 0), np.where(has, first_max, 0))
    # no outliers: background first frame, blend mask 0
    none = ~has
    assert np.all(mask[none] == 0)
    assert np.array_equal(img[none], x[0][none])


def test_outlier_first_last_extreme_select_the_right_frame():
    rng = np.random.default_rng(3)
    x = _stack(rng, 30, 5, 6, 3, 8, outlier_frac=0.3)
    thr = O.threshold(True, 0.05, 0.05)  # max == min: any outlier blends fully
    for mode in (0, 1, 2):
        img, mask, _, st = O.outlier(x, thr, 3, mode, stats=True)
        med = np.median(x.astype(np.float64), axis=0).astype(np.float32)
        d = _numpy_distances(x, med, np.zeros_like(med), (1, 1, 1, 1), True)
        out = d >= F32(thr.min) * F32(thr.min)
        for yy in range(5):
            for xx in range(6):
                ts = np.nonzero(out[:, yy, xx])[0]
                if len(ts) == 0:
                    assert np.array_equal(img[yy, xx], rround(med[yy, xx]).astype(np.uint8))
                    continue
                t = {0: ts[0], 1: ts[-1], 2: ts[np.argmax(d[ts, yy, xx])]}[mode]
                assert np.array_equal(img[yy, xx], x[t, yy, xx]), (mode, yy, xx)
                assert mask[yy, xx, 0] == 255


def test_forward_backward_with_full_blend_end_on_last_and_first_outlier():
    rng = np.random.default_rng(5)
    x = _stack(rng, 25, 4, 5, 3, 8, outlier_frac=0.4)
    thr = O.threshold(True, 0.05, 0.05)
    f, _, _, st = O.outlier(x, thr, 0, 4, stats=True)
    b, _, _, _ = O.outlier(x, thr, 0, 5, stats=True)
    l, _, _ = O.outlier(x, thr, 0, 1)
    fi, _, _ = O.outlier(x, thr, 0, 0)
    many = st["num_outliers"] >= 1
    assert np.array_equal(f[many], l[many])
    assert np.array_equal(b[many], fi[many])


def test_background_average_without_outliers_is_rounded_mean():
    rng = np.random.default_rng(9)
    x = rng.integers(100, 110, (9, 4, 4, 3)).astype(np.uint8)  # |diff| < 12.75: no outliers
    img, mask, _ = O.outlier(x, O.threshold(True, 0.05, 0.2), 2, 2)
    mean = (x.astype(np.float32).sum(axis=0) / F32(9)).astype(np.float32)
    assert np.array_equal(img, rround(mean).astype(np.uint8))
    assert np.all(mask[..., :3] == 0)


def _round_half_away(num, den):
    """Exact half-away-from-zero rounding of num/den for non-negative integers."""
    return (2 * num + den) // (2 * den)


def _outlier_sets(x, thr_min):
    med = np.median(x.astype(np.float64), axis=0).astype(np.float32)
    d = _numpy_distances(x, med, np.zeros_like(med), (1, 1, 1, 1), True)
    return d >= F32(thr_min) * F32(thr_min)


ZERO_FADE = dict(repeat=False, absolute=True, frames=[0, 1], values=[0.0, 0.0])


def test_background_average_with_outliers_closed_form():
    """A fade that is 0 on every frame makes every blend value 0 (src/color.rs:5: blend <= 0
    leaves the pixel), so the output is the background itself.  With k outliers of N samples,
    sum S and outlier sum O, src/chrono.rs:306-331 evaluates mean * N/(N-k) - O/N, i.e.
    S/(N-k) - O/N (not the mean of the other samples; the source marks the equation "TODO:
    check"), rounded half away and saturated to u8.  The exact rational of that expression
    has denominator N (N-k), so it sits at least 1/(2N(N-k)) from a half unless it is one;
    for u8 and N = 25 the f32 evaluation errs by < 255 N 2^-22, well inside that gap, so only
    exact halves (skipped) could round either way."""
    rng = np.random.default_rng(21)
    thr = O.threshold(True, 0.05, 0.2)
    checked, ks = 0, []
    for frac in (0.3, 0.04):  # many-outlier pixels, and single-outlier ones (their own branch)
        x = _stack(rng, 25, 6, 7, 3, 8, outlier_frac=frac)
        img, mask, _, st = O.outlier(x, thr, 2, 2, fade=O.Fade(**ZERO_FADE), stats=True)
        out = _outlier_sets(x, thr.min)
        assert np.array_equal(st["num_outliers"], out.sum(axis=0))
        assert np.all(mask == 0)
        n = x.shape[0]
        for yy in range(x.shape[1]):
            for xx in range(x.shape[2]):
                k = int(out[:, yy, xx].sum())
                if k == n:
                    continue
                ks.append(k)
                for ch in range(3):
                    tot = int(x[:, yy, xx, ch].astype(np.int64).sum())
                    osum = int(x[out[:, yy, xx], yy, xx, ch].astype(np.int64).sum())
                    num, den = tot * n - osum * (n - k), n * (n - k)
                    if (2 * num) % den == 0 and (2 * num // den) % 2 != 0:
                        continue  # exact half
                    want = 0 if num < 0 else min(255, _round_half_away(num, den))
                    assert img[yy, xx, ch] == want, (frac, yy, xx, ch, k)
                    checked += 1
    assert checked > 200 and ks.count(0) > 0 and ks.count(1) > 3 and sum(k > 1 for k in ks) > 5


def test_average_outlier_mode_with_full_blend_is_rounded_outlier_mean():
    """src/chrono.rs:430-466: with two or more outliers the blended sample is the rounded f32
    mean of the outliers' samples; min == max makes every blend value 1, which copies it
    (src/color.rs:6-9).  4-channel stacks: the reference blends its 4-entry mean into the
    pixel, which only fits RGBA at blend 1.  For u8 the f32 quotient of exact sums rounds like
    the exact rational (halves included: they are exact in f32)."""
    rng = np.random.default_rng(23)
    x = _stack(rng, 30, 5, 6, 4, 8, outlier_frac=0.35)
    thr = O.threshold(True, 0.05, 0.05)
    img, mask, _, st = O.outlier(x, thr, 0, 3, stats=True)
    out = _outlier_sets(x, thr.min)
    assert np.array_equal(st["num_outliers"], out.sum(axis=0))
    multi = 0
    for yy in range(x.shape[1]):
        for xx in range(x.shape[2]):
            ts = np.nonzero(out[:, yy, xx])[0]
            if len(ts) == 0:
                assert np.array_equal(img[yy, xx], x[0, yy, xx]) and np.all(mask[yy, xx] == 0)
                continue
            assert np.all(mask[yy, xx] == 255)
            if len(ts) == 1:
                assert np.array_equal(img[yy, xx], x[ts[0], yy, xx])
                continue
            multi += 1
            for ch in range(4):
                osum = int(x[ts, yy, xx, ch].astype(np.int64).sum())
                assert img[yy, xx, ch] == _round_half_away(osum, len(ts)), (yy, xx, ch)
    assert multi > 5


def test_partial_blends_with_dyadic_fades_are_exact_convex_steps():
    """src/chrono.rs:379-428 with src/color.rs:4-44.  Every outlier blends fully (min = max)
    and the fade gives each frame a blend of 1/4, 1/2 or 3/4, so each step a + (b - a) t of
    u8 values is exact in f32 while the value fits 24 bits (up to 8 outliers): forward and
    backward must equal the exact rational sequence of convex steps in frame order / reverse
    order, rounded half away, and the mask round((1 - prod(1 - t)) * 255); one outlier must
    give round(bg + (s - bg) t) and round(t * 255).  Background first."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(29)
    n = 20
    x = _stack(rng, n, 6, 7, 3, 8, outlier_frac=0.2)
    tvals = rng.choice([0.25, 0.5, 0.75], n)
    fade = O.Fade(False, True, list(range(n)), [float(t) for t in tvals])
    thr = O.threshold(True, 0.05, 0.05)
    out = _outlier_sets(x, thr.min)
    seen = {1: 0, 2: 0}
    for mode in (4, 5):
        img, mask, _, st = O.outlier(x, thr, 0, mode, fade=fade, stats=True)
        assert np.array_equal(st["num_outliers"], out.sum(axis=0))
        for yy in range(x.shape[1]):
            for xx in range(x.shape[2]):
                ts = list(np.nonzero(out[:, yy, xx])[0])
                k = len(ts)
                if k == 0 or k == n or k > 8:
                    continue
                bg = x[[t for t in range(n) if t not in ts][0], yy, xx]
                order = ts if mode == 4 else ts[::-1]
                inv = Fr(1)
                for ch in range(3):
                    a = Fr(int(bg[ch]))
                    for t in order:
                        a = a + (int(x[t, yy, xx, ch]) - a) * Fr(float(tvals[t]))
                    want = _round_half_away(a.numerator, a.denominator)
                    assert img[yy, xx, ch] == want, (mode, yy, xx, ch, k)
                for t in order:
                    inv *= 1 - Fr(float(tvals[t]))
                m = (1 - inv) * 255
                assert np.all(mask[yy, xx] == _round_half_away(m.numerator, m.denominator)), (mode, yy, xx)
                seen[min(k, 2)] += 1
    assert seen[1] > 3 and seen[2] > 10


def test_relative_fade_reads_the_table_from_the_last_frame():
    """src/chrono.rs:496-502: a relative fade takes Fade::get(N - t - 1) for frame t, so the
    table runs backwards from the last frame.  Every outlier blends fully (min = max) and the
    table gives frame t the dyadic value tvals[N - t - 1], so forward / backward must equal
    the exact rational convex steps with those values, and a single outlier
    round(bg + (s - bg) t) with t = tvals[N - t - 1].  An index of t or N - t would pick other
    table entries and fail."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(41)
    n = 18
    x = _stack(rng, n, 6, 7, 3, 8, outlier_frac=0.2)
    tvals = rng.choice([0.25, 0.5, 0.75], n)
    assert len(set(tvals[:6])) > 1
    fade = O.Fade(False, False, list(range(n)), [float(t) for t in tvals])
    thr = O.threshold(True, 0.05, 0.05)
    out = _outlier_sets(x, thr.min)
    seen = 0
    for mode in (4, 5):
        img, mask, _ = O.outlier(x, thr, 0, mode, fade=fade)
        for yy in range(x.shape[1]):
            for xx in range(x.shape[2]):
                ts = list(np.nonzero(out[:, yy, xx])[0])
                if not ts or len(ts) == n or len(ts) > 8:
                    continue
                bg = x[[t for t in range(n) if t not in ts][0], yy, xx]
                order = ts if mode == 4 else ts[::-1]
                inv = Fr(1)
                for t in order:
                    inv *= 1 - Fr(float(tvals[n - t - 1]))
                for ch in range(3):
                    a = Fr(int(bg[ch]))
                    for t in order:
                        a = a + (int(x[t, yy, xx, ch]) - a) * Fr(float(tvals[n - t - 1]))
                    assert img[yy, xx, ch] == _round_half_away(a.numerator, a.denominator), (mode, yy, xx, ch)
                m = (1 - inv) * 255
                assert np.all(mask[yy, xx] == _round_half_away(m.numerator, m.denominator)), (mode, yy, xx)
                seen += 1
    assert seen > 20


@pytest.mark.parametrize("repeat", [False, True])
def test_simple_relative_fade_by_hand(repeat):
    """src/simple.rs:170-176 with Fade::get (src/options.rs:113-140): darker mode over seven
    frames that each set a new extreme blends frame t with get(N - t - 1).  The table has three
    entries (keys 0, 1, 2 -> 1/4, 1/2, 3/4), so indices 6..0 clamp to the last entry or wrap
    modulo 3; the expected pixel is the sequence of rounded blends written out by hand."""
    x = np.zeros((7, 1, 1, 1), np.uint8)
    x[:, 0, 0, 0] = [240, 200, 160, 120, 80, 40, 0]
    fade = O.Fade(repeat, False, [0, 1, 2], [0.25, 0.5, 0.75])
    # index N - t - 1 = 6, 5, 4, 3, 2, 1, 0 for t = 0..6
    table = {False: [0.75, 0.75, 0.75, 0.75, 0.75, 0.5, 0.25],   # clamp: indices >= 3 take entry 2
             True: [0.25, 0.75, 0.5, 0.25, 0.75, 0.5, 0.25]}[repeat]  # repeat: index mod 3
    o = 0
    for t, v in enumerate(x[:, 0, 0, 0]):
        assert fade.get(6 - t) == F32(table[t])
        o = int(rround(F32(o) + (F32(int(v)) - F32(o)) * F32(table[t])))
    img = O.simple(x, (1, 1, 1, 1), fade=fade, darker=True)
    assert img[0, 0, 0] == o


def test_average_mode_partial_blend_uses_frame_zero_fade():
    """src/chrono.rs:429-481: with two or more outliers the average mode blends the rounded
    outlier mean with the fade of frame 0 (the tuple's index is 0), with one outlier that
    outlier with its own frame's fade; full blend values (min = max) and dyadic fades make
    round(bg + (s - bg) t) and round(t * 255) exact.  A partial blend zips only C entries,
    so RGB is fine here."""
    rng = np.random.default_rng(31)
    n = 20
    tvals = rng.choice([0.25, 0.5, 0.75], n)
    fade = O.Fade(False, True, list(range(n)), [float(t) for t in tvals])
    thr = O.threshold(True, 0.05, 0.05)
    seen = {1: 0, 2: 0}
    for frac in (0.25, 0.04):
        x = _stack(rng, n, 6, 7, 3, 8, outlier_frac=frac)
        out = _outlier_sets(x, thr.min)
        img, mask, _, st = O.outlier(x, thr, 0, 3, fade=fade, stats=True)
        for yy in range(x.shape[1]):
            for xx in range(x.shape[2]):
                ts = list(np.nonzero(out[:, yy, xx])[0])
                k = len(ts)
                if k == 0 or k == n:
                    continue
                bg = x[[t for t in range(n) if t not in ts][0], yy, xx].astype(np.int64)
                t = tvals[ts[0]] if k == 1 else tvals[0]
                for ch in range(3):
                    sm = int(x[ts[0], yy, xx, ch]) if k == 1 else _round_half_away(
                        int(x[ts, yy, xx, ch].astype(np.int64).sum()), k)
                    v = bg[ch] * 4 + (sm - bg[ch]) * int(t * 4)  # 4 * (bg + (s - bg) t)
                    assert img[yy, xx, ch] == _round_half_away(v, 4), (frac, yy, xx, ch, k)
                assert np.all(mask[yy, xx] == _round_half_away(int(t * 4) * 255, 4))
                seen[min(k, 2)] += 1
    assert seen[1] > 3 and seen[2] > 10


def test_random_background_draws_uniformly_from_the_candidates():
    """src/chrono.rs:344-349, 530-555 with every blend value 0 (the output is the background
    frame).  The reference draws from an OS-seeded ThreadRng, so single values cannot be
    pinned; what is fixed: no outliers -> gen_range(0, N) over every frame; outliers {1, 4}
    of N = 6 -> the swap loop leaves positions [0, 4) holding frames {0, 5, 2, 3} (swap(1, 5),
    then swap(4, 4)), and the draw is uniform over them.  Many pixels with one pattern check
    the support exactly and the frequencies to within 25%."""
    n, h, w = 6, 40, 40
    x = np.full((n, h, w, 1), 100, np.uint8)
    for t in range(n):
        x[t, ..., 0] += t  # distinct per frame, all within the threshold
    img, mask, _, st = O.outlier(x, O.threshold(True, 0.05, 0.2), 1, 2, fade=O.Fade(**ZERO_FADE),
                                 stats=True, seed=7)
    assert np.all(st["num_outliers"] == 0)
    bg = st["bg_index"].reshape(-1)
    assert np.array_equal(img[..., 0].reshape(-1), 100 + bg)
    counts = np.bincount(bg, minlength=n)
    assert len(counts) == n and np.all(np.abs(counts - h * w / n) < 0.25 * h * w / n)
    x[1, ..., 0] = 250
    x[4, ..., 0] = 0
    img, mask, _, st = O.outlier(x, O.threshold(True, 0.05, 0.2), 1, 2, fade=O.Fade(**ZERO_FADE),
                                 stats=True, seed=7)
    assert np.all(st["num_outliers"] == 2)
    bg = st["bg_index"].reshape(-1)
    assert np.array_equal(img[..., 0].reshape(-1), x[bg, 0, 0, 0])
    counts = np.bincount(bg, minlength=n)
    assert set(np.nonzero(counts)[0].tolist()) == {0, 2, 3, 5}
    assert np.all(np.abs(counts[[0, 2, 3, 5]] - h * w / 4) < 0.25 * h * w / 4)


def test_background_first_and_warning_for_all_outlier_pixel():
    x = np.zeros((4, 1, 1, 3), np.uint8)
    x[:, 0, 0, 0] = [0, 0, 255, 255]  # median 127.5: every frame is 127.5 away
    img, mask, warn, st = O.outlier(x, O.threshold(True, 0.05, 0.2), 0, 2, stats=True)
    assert st["num_outliers"][0, 0] == 4 and warn == 1 and st["warning"][0, 0] == 1


def test_rel_threshold_needs_three_samples():
    x = np.zeros((2, 1, 1, 3), np.uint8)
    with pytest.raises(ValueError):
        O.outlier(x, O.threshold(False, 2.0, 4.0), 0, 2)


def test_random_background_swap_candidates_can_hold_outliers():
    # src/chrono.rs:544-554: N=3, outliers {0, 2}: swaps (0,2) then (2,1) leave [2, 0, 1];
    # the single candidate is index 2, an outlier (reference quirk).
    x = np.zeros((3, 1, 1, 1), np.uint8)
    x[:, 0, 0, 0] = [0, 100, 200]  # median 100: frames 0 and 2 are outliers
    img, mask, warn, st = O.outlier(x, O.threshold(True, 0.05, 0.2), 1, 2, stats=True, seed=123)
    assert st["num_outliers"][0, 0] == 2
    assert st["bg_index"][0, 0] == 2


def test_sample_subset_median():
    rng = np.random.default_rng(1)
    x = _stack(rng, 20, 3, 3, 3, 8)
    pos = np.array([1, 4, 5, 9, 12, 17, 19], np.uint32)
    img, mask, warn, st = O.outlier(x, O.threshold(True, 0.05, 0.2), 0, 2, sample_pos=pos, stats=True)
    med = np.median(x[pos].astype(np.float64), axis=0).astype(np.float32)
    assert np.array_equal(st["median"][..., :3], med)


# --------------------------------------------------------------- darker/lighter

@pytest.mark.parametrize("darker", [True, False])
def test_simple_without_fade_is_first_argextreme(darker):
    rng = np.random.default_rng(2 + darker)
    x = rng.integers(0, 4, (17, 5, 6, 3)).astype(np.uint8)  # many ties
    w = (1.0, 1.0, 1.0, 1.0)
    img = O.simple(x, w, darker=darker)
    v = x.astype(np.float32).sum(axis=-1)
    idx = np.argmin(v, axis=0) if darker else np.argmax(v, axis=0)
    exp = np.take_along_axis(x, idx[None, :, :, None], axis=0)[0]
    assert np.array_equal(img, exp)


def test_simple_with_fade_blends_on_each_new_extreme():
    x = np.zeros((3, 1, 1, 1), np.uint8)
    x[:, 0, 0, 0] = [200, 100, 50]  # each frame is a new darker extreme
    fade = O.Fade(False, True, [0, 2], [0.5, 0.5])
    img = O.simple(x, (1, 1, 1, 1), fade=fade, darker=True)
    o = 0
    for v in (200, 100, 50):
        o = int(rround(F32(o) + (F32(v) - F32(o)) * F32(0.5)))
    assert img[0, 0, 0] == o


# ------------------------------------------------------------------- shake + crop

def _textured_frames(n, h, w, shifts, seed=0):
    rng = np.random.default_rng(seed)
    world = rng.integers(0, 256, (h + 40, w + 40, 3)).astype(np.uint8)
    out = np.zeros((n, h, w, 3), np.uint8)
    for t, (dx, dy) in enumerate(shifts):
        out[t] = world[20 - dy:20 - dy + h, 20 - dx:20 - dx + w]
    return out


def test_shake_recovers_translations_and_matches_brute_force():
    shifts = [(0, 0), (2, -1), (-3, 3), (1, 0), (0, -2)]
    x = _textured_frames(5, 40, 48, shifts)
    anchors = [(16, 14), (30, 24)]
    off = O.shake(x, anchors, 4, 4)
    assert [tuple(o) for o in off] == shifts
    # brute force int64 SSD, first minimum in row-major order
    r, s = 4, 4
    for t in range(1, 5):
        best, bi = None, None
        for oy in range(2 * s + 1):
            for ox in range(2 * s + 1):
                tot = 0
                for cx, cy in anchors:
                    win = x[0, cy - r:cy + r + 1, cx - r:cx + r + 1].astype(np.int64)
                    yy, xx = cy + oy - s, cx + ox - s
                    img = x[t, yy - r:yy + r + 1, xx - r:xx + r + 1].astype(np.int64)
                    tot += int(((win - img) ** 2).sum())
                if best is None or tot < best:
                    best, bi = tot, (ox - s, oy - s)
        assert tuple(off[t]) == bi


def test_shake_u8_sum_wraps_like_a_release_build_i32():
    """src/shake.rs:350-379 sums squared differences of u8 samples in an i32 that wraps in the
    reference's release build (Cargo.toml: overflow-checks = false).  One anchor of radius 52
    (105 x 105 x 3 = 33,075 terms) against a frame of 255s gives 33,075 * 255^2 =
    2,150,701,875 > 2^31 for the candidates that miss a column of zeros and 32,760 * 255^2 <
    2^31 for those that contain it, so the exact arg-min (first minimum over (oy, ox) in
    row-major order) and the arg-min of the sums read modulo 2^32 as i32 differ; the oracle
    must return the wrapped one."""
    r, s = 52, 1
    h = w = 120
    cx = cy = 60
    x = np.zeros((2, h, w, 3), np.uint8)
    x[1] = 255
    x[1, :, cx - r - s, :] = 0  # only windows shifted by ox = -1 reach this column
    exact, wrapped = {}, {}
    for oy in range(2 * s + 1):
        for ox in range(2 * s + 1):
            win = x[0, cy - r:cy + r + 1, cx - r:cx + r + 1].astype(np.int64)
            yy, xx = cy + oy - s, cx + ox - s
            img = x[1, yy - r:yy + r + 1, xx - r:xx + r + 1].astype(np.int64)
            tot = int(((win - img) ** 2).sum())
            exact[(oy, ox)] = tot
            wrapped[(oy, ox)] = (tot + (1 << 31)) % (1 << 32) - (1 << 31)
    order = [(oy, ox) for oy in range(2 * s + 1) for ox in range(2 * s + 1)]
    ex = min(order, key=lambda k: (exact[k], order.index(k)))
    wr = min(order, key=lambda k: (wrapped[k], order.index(k)))
    assert max(exact.values()) > (1 << 31) > min(exact.values())
    assert ex != wr
    off = O.shake(x, [(cx, cy)], r, s)
    assert tuple(off[1]) == (wr[1] - s, wr[0] - s)
    assert tuple(off[1]) != (ex[1] - s, ex[0] - s)


def test_shake_anchor_out_of_range():
    x = np.zeros((2, 20, 20, 3), np.uint8)
    with pytest.raises(ValueError):
        O.shake(x, [(5, 5)], 3, 3)


def test_crop_create_closed_form():
    off = [(0, 0), (2, -1), (-3, 3)]
    crop, w, h = O.crop_create(off, 100, 80)
    assert (w, h) == (100 - 3 - 2, 80 - 1 - 3)
    assert crop.tolist() == [[3, 1], [5, 0], [0, 4]]
    assert O.crop_create([(0, 0), (0, 0)], 10, 10) is None


# -------------------------------------------------------------------- video windows

def test_video_windows_full_windows():
    # --video-in 0/60/. --video-out 0/540/. over 600 frames: 540 windows of 60 frames
    wins = O.video_windows((0, 60, 1), (0, 540, 1), 600)
    assert len(wins) == 540
    assert all(len(ix) == 60 for _, ix in wins)
    assert wins[0] == (0, list(range(0, 60)))
    assert wins[539] == (539, list(range(539, 599)))


def test_video_windows_default_out_ramps_and_never_uses_last_frame():
    # --video-in 0/60/. alone: v_lower = -59, 659 windows ramping 1..60..1 frames;
    # the end clamp min(N + (f+e)%1 - 1, f+e) keeps frame 599 out of every window
    wins = O.video_windows((0, 60, 1), (None, None, 1), 600)
    assert len(wins) == 659
    sizes = [len(ix) for _, ix in wins]
    assert sizes[0] == 1 and max(sizes) == 60 and sizes[-1] == 0
    assert all(599 not in ix for _, ix in wins)


def test_video_windows_step():
    wins = O.video_windows((0, 10, 3), (0, 4, 2), 20)
    # frames 0 and 2: start = max(st % 3, f), end = min(20 + (f+10)%3 - 3, f + 10)
    assert wins == [(0, [0, 3, 6, 9]), (2, [2, 5, 8, 11])]
