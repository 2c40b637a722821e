"""Writes the golden fixtures of this directory: small seeded stacks and the oracle's outputs
for them.  Run from the repository root after building the oracle (make -C oracle):

    python tests/golden/make_golden.py

Contents (one compressed .npz per family; `cfg` holds the option tuple of each case):
  outlier_rgb8.npz   16 frames of 32x24 RGB8 (the 640x480 RGB8 workload's generator: gradient,
                     noise, a moving dark disk); threshold abs 0.05/0.2 and rel 2/4 x outlier
                     first|last|extreme|average|forward|backward x background first|median|average
                     x fade none|clamp/abs/0,0.3/15,1.0 -> image + blend mask; per threshold the
                     per-pixel median, inverse IQR, outlier count and extreme frame.
  outlier_rgb16.npz  20 frames of 32x24 RGB16 with per-frame crop origins (the 4K RGB16
                     workload's generator at tile size): abs 0.05/0.2, forward and extreme,
                     background first.
  simple_rgb8.npz    darker / lighter with and without fade on the RGB8 stack.
  shake_rgb8.npz     8 translated frames of 64x48 RGB8 with a static texture: shake offsets for
                     two anchors, r = 4, s = 3.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "chrono-photo_amd")]

import oracle_py as O  # noqa: E402
import chrono_photo_amd as cpa  # noqa: E402

THRESHOLDS = [("abs", True, 0.05, 0.2), ("rel", False, 2.0, 4.0)]
OUTLIER = ["first", "last", "extreme", "average", "forward", "backward"]
BACKGROUND = {"first": 0, "median": 3, "average": 2}
FADE = (False, True, [0, 15], [0.3, 1.0])  # clamp/abs/0,0.3/15,1.0


def rgb8_stack():
    sc = cpa.Scene(seed=11, width=32, height=24, channels=3, bytes_per_sample=1, n_frames=16,
                   base_lo=40, base_hi=200, sigma=3.0, n_disks=1, disk_r_min=6, disk_r_max=6,
                   disk_val_lo=15, disk_val_hi=25, blob_permille=150, blob_r=5, blob_val=230)
    return sc.generate()


def rgb16_stack():
    sc = cpa.Scene(seed=12, width=40, height=32, channels=3, bytes_per_sample=2, n_frames=20,
                   base_lo=8000, base_hi=50000, sigma=200.0, n_disks=2, disk_r_min=4, disk_r_max=6,
                   disk_val_lo=2000, disk_val_hi=4000, shake_max=3, texture_amp=3000)
    return sc.generate(), [sc.native()]


def outlier_rgb8(path):
    x = rgb8_stack()
    out = {"frames": x}
    cfg = []
    for tname, absolute, lo, hi in THRESHOLDS:
        thr = O.threshold(absolute, lo, hi, 1)
        _, _, _, st = O.outlier(x, thr, 0, 2, stats=True)
        out[f"median_{tname}"] = st["median"][..., :3].copy()
        out[f"iqr_inv_{tname}"] = st["iqr_inv"][..., :3].copy()
        out[f"num_{tname}"] = st["num_outliers"].copy()
        out[f"max_index_{tname}"] = st["max_index"].copy()
        for oi, oname in enumerate(OUTLIER):
            for bname, bg in BACKGROUND.items():
                for fade_on in (False, True):
                    fade = O.Fade(*FADE) if fade_on else O.Fade.none()
                    img, mask, warn = O.outlier(x, thr, bg, oi, fade=fade)
                    key = f"{tname}_{oname}_{bname}_{'fade' if fade_on else 'nofade'}"
                    out["img_" + key] = img
                    out["mask_" + key] = mask
                    cfg.append(key)
    out["cfg"] = np.array(cfg)
    np.savez_compressed(path, **out)


def outlier_rgb16(path):
    x, _ = rgb16_stack()
    n = x.shape[0]
    rng = np.random.default_rng(5)
    off = rng.integers(-3, 4, (n, 2)).astype(np.int32)
    off[0] = 0
    crop = O.crop_create(off, x.shape[2], x.shape[1])
    crop_xy, ow, oh = crop
    out = {"frames": x, "crop_xy": crop_xy, "out_wh": np.array([ow, oh], np.int32)}
    thr = O.threshold(True, 0.05, 0.2, 2)
    cfg = []
    for oi in (2, 4):
        img, mask, warn, st = O.outlier(x, thr, 0, oi, crop_xy=crop_xy, out_w=ow, out_h=oh, stats=True)
        key = OUTLIER[oi]
        out["img_" + key] = img
        out["mask_" + key] = mask
        out["median_" + key] = st["median"][..., :3].copy()
        out["num_" + key] = st["num_outliers"].copy()
        cfg.append(key)
    out["cfg"] = np.array(cfg)
    np.savez_compressed(path, **out)


def simple_rgb8(path):
    x = rgb8_stack()
    out = {"frames": x}
    w = (0.5, 1.0, 0.25, 1.0)
    for darker in (True, False):
        name = "darker" if darker else "lighter"
        out[name] = O.simple(x, weights=w, darker=darker)
        out[name + "_fade"] = O.simple(x, weights=w, fade=O.Fade(*FADE), darker=darker)
    out["weights"] = np.array(w, np.float32)
    np.savez_compressed(path, **out)


def shake_rgb8(path):
    sc = cpa.Scene(seed=13, width=64, height=48, channels=3, bytes_per_sample=1, n_frames=8,
                   base_lo=60, base_hi=180, sigma=2.0, n_disks=0, shake_max=3, texture_amp=50)
    x = sc.generate()
    anchors = np.array([[20, 16], [44, 30]], np.int32)
    off = O.shake(x, anchors, 4, 3)
    np.savez_compressed(path, frames=x, anchors=anchors, r=np.int32(4), s=np.int32(3), offsets=off)


def main():
    O.build()
    outlier_rgb8(os.path.join(HERE, "outlier_rgb8.npz"))
    outlier_rgb16(os.path.join(HERE, "outlier_rgb16.npz"))
    simple_rgb8(os.path.join(HERE, "simple_rgb8.npz"))
    shake_rgb8(os.path.join(HERE, "shake_rgb8.npz"))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
