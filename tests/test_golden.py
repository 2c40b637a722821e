"""Committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them on the CPU, and the HIP path must match them bit for
bit on the GPU."""
import os

import numpy as np
import pytest

import chrono_photo_amd as cpa
import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUTLIER = ["first", "last", "extreme", "average", "forward", "backward"]
BACKGROUND = {"first": 0, "median": 3, "average": 2}
THRESHOLDS = {"abs": (True, 0.05, 0.2), "rel": (False, 2.0, 4.0)}
FADE = (False, True, [0, 15], [0.3, 1.0])


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def parse(key):
    tname, oname, bname, fname = key.split("_")
    return tname, OUTLIER.index(oname), BACKGROUND[bname], fname == "fade"


# ------------------------------------------------------------------ CPU: the oracle
def test_oracle_reproduces_outlier_rgb8():
    z = load("outlier_rgb8.npz")
    x = z["frames"]
    for tname, (absolute, lo, hi) in THRESHOLDS.items():
        _, _, _, st = O.outlier(x, O.threshold(absolute, lo, hi, 1), 0, 2, stats=True)
        assert np.array_equal(st["median"][..., :3], z["median_" + tname])
        assert np.array_equal(st["iqr_inv"][..., :3], z["iqr_inv_" + tname])
        assert np.array_equal(st["num_outliers"], z["num_" + tname])
        assert np.array_equal(st["max_index"], z["max_index_" + tname])
        assert (z["num_" + tname] > 0).any() and (z["num_" + tname] == 0).any()
    for key in z["cfg"]:
        tname, oi, bg, fade_on = parse(str(key))
        absolute, lo, hi = THRESHOLDS[tname]
        fade = O.Fade(*FADE) if fade_on else O.Fade.none()
        img, mask, _ = O.outlier(x, O.threshold(absolute, lo, hi, 1), bg, oi, fade=fade)
        assert np.array_equal(img, z["img_" + key]), key
        assert np.array_equal(mask, z["mask_" + key]), key


def test_oracle_reproduces_simple_and_shake():
    z = load("simple_rgb8.npz")
    for darker in (True, False):
        name = "darker" if darker else "lighter"
        assert np.array_equal(O.simple(z["frames"], weights=z["weights"], darker=darker), z[name])
        assert np.array_equal(O.simple(z["frames"], weights=z["weights"], fade=O.Fade(*FADE), darker=darker),
                              z[name + "_fade"])
    s = load("shake_rgb8.npz")
    off = O.shake(s["frames"], s["anchors"], int(s["r"]), int(s["s"]))
    assert np.array_equal(off, s["offsets"])
    assert (s["offsets"] != 0).any()


def test_shake_fixture_is_the_scene_translation():
    """The fixture's offsets are the generator's known per-frame translations (a pin that does
    not go through the oracle)."""
    s = load("shake_rgb8.npz")
    sc = cpa.Scene(seed=13, width=64, height=48, channels=3, bytes_per_sample=1, n_frames=8,
                   base_lo=60, base_hi=180, sigma=2.0, n_disks=0, shake_max=3, texture_amp=50)
    assert np.array_equal(sc.generate(), s["frames"])
    truth = sc.shake_truth()
    assert np.array_equal(s["offsets"], truth)


# ------------------------------------------------------------------ GPU: the HIP path
@pytest.mark.gpu
def test_hip_matches_outlier_rgb8(ctx):
    z = load("outlier_rgb8.npz")
    st = cpa.Stack.from_array(ctx, z["frames"])
    for tname, (absolute, lo, hi) in THRESHOLDS.items():
        proc = cpa.OutlierProcessor(cpa.Threshold(absolute, lo, hi), 0, 2)
        _, _, stats = proc.process(st, with_stats=True)
        assert np.array_equal(stats["median"][..., :3], z["median_" + tname]), tname
        assert np.array_equal(stats["iqr_inv"][..., :3], z["iqr_inv_" + tname]), tname
        assert np.array_equal(stats["num_outliers"], z["num_" + tname]), tname
        assert np.array_equal(stats["max_index"], z["max_index_" + tname]), tname
    for key in z["cfg"]:
        key = str(key)
        tname, oi, bg, fade_on = parse(key)
        absolute, lo, hi = THRESHOLDS[tname]
        fade = cpa.Fade(False, True, list(zip(FADE[2], FADE[3]))) if fade_on else cpa.Fade.none()
        proc = cpa.OutlierProcessor(cpa.Threshold(absolute, lo, hi), bg, oi, fade=fade)
        img, mask = proc.process(st)
        assert np.array_equal(img, z["img_" + key]), key
        assert np.array_equal(mask, z["mask_" + key]), key
    st.close()


@pytest.mark.gpu
def test_hip_matches_outlier_rgb16_cropped(ctx):
    z = load("outlier_rgb16.npz")
    st = cpa.Stack.from_array(ctx, z["frames"])
    ow, oh = (int(v) for v in z["out_wh"])
    for key in z["cfg"]:
        key = str(key)
        proc = cpa.OutlierProcessor(cpa.Threshold(True, 0.05, 0.2), 0, OUTLIER.index(key))
        img, mask, stats = proc.process(st, crop=(z["crop_xy"], ow, oh), with_stats=True)
        assert np.array_equal(img, z["img_" + key]), key
        assert np.array_equal(mask, z["mask_" + key]), key
        assert np.array_equal(stats["median"][..., :3], z["median_" + key]), key
        assert np.array_equal(stats["num_outliers"], z["num_" + key]), key
    st.close()


@pytest.mark.gpu
def test_hip_matches_simple_and_shake(ctx):
    z = load("simple_rgb8.npz")
    st = cpa.Stack.from_array(ctx, z["frames"])
    fade = cpa.Fade(False, True, list(zip(FADE[2], FADE[3])))
    for darker in (True, False):
        name = "darker" if darker else "lighter"
        assert np.array_equal(cpa.SimpleProcessor(z["weights"], None, darker).process(st), z[name])
        assert np.array_equal(cpa.SimpleProcessor(z["weights"], fade, darker).process(st), z[name + "_fade"])
    st.close()
    s = load("shake_rgb8.npz")
    st = cpa.Stack.from_array(ctx, s["frames"])
    off = cpa.ShakeAnalyzer().analyze(st, s["anchors"], int(s["r"]), int(s["s"]))
    assert np.array_equal(off, s["offsets"])
    st.close()
