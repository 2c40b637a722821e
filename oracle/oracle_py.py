"""ctypes binding of the CPU oracle (oracle/chrono_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, smoke() and bench.py's cpu_baseline leg use this module
as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libchrono_oracle.so")


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


class or_threshold(C.Structure):
    _fields_ = [("absolute", C.c_int32), ("min", C.c_float), ("max", C.c_float), ("scale", C.c_float)]


class or_fade(C.Structure):
    _fields_ = [("is_none", C.c_int32), ("repeat", C.c_int32), ("absolute", C.c_int32),
                ("offset", C.c_int32), ("len", C.c_uint32), ("values", C.POINTER(C.c_float))]


class or_outlier_params(C.Structure):
    _fields_ = [("threshold", or_threshold), ("background", C.c_int32), ("outlier", C.c_int32),
                ("weights", C.c_float * 4), ("fade", or_fade), ("seed", C.c_uint64)]


class or_view(C.Structure):
    _fields_ = [("data", C.c_void_p), ("bytes_per_sample", C.c_uint32), ("channels", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("slots", C.POINTER(C.c_uint32)),
                ("n", C.c_uint32), ("crop_xy", C.POINTER(C.c_int32)), ("out_w", C.c_uint32),
                ("out_h", C.c_uint32)]


STATS_DTYPE = np.dtype([("median", "<f4", (4,)), ("iqr_inv", "<f4", (4,)), ("num_outliers", "<u4"),
                        ("max_index", "<u4"), ("max_dist_sq", "<f4"), ("bg_index", "<u4"),
                        ("warning", "<u4"), ("pad", "<u4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        L.or_threshold_new.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_float, P(or_threshold)]
        L.or_blend_value.argtypes = [P(or_threshold), C.c_float]
        L.or_blend_value.restype = C.c_float
        L.or_fade_table.argtypes = [P(C.c_int32), P(C.c_float), C.c_uint32, P(C.c_float), C.c_uint32]
        L.or_fade_table.restype = C.c_uint32
        L.or_fade_get.argtypes = [P(or_fade), C.c_int32]
        L.or_fade_get.restype = C.c_float
        L.or_fade_pick.argtypes = [P(or_fade), C.c_int32, C.c_int32, C.c_int32]
        L.or_fade_pick.restype = C.c_float
        L.or_median.argtypes = [P(C.c_uint32), C.c_uint32]
        L.or_median.restype = C.c_float
        L.or_quantile.argtypes = [P(C.c_uint32), C.c_uint32, C.c_float]
        L.or_quantile.restype = C.c_float
        L.or_blend_into_u8.argtypes = [P(C.c_uint8), P(C.c_uint8), C.c_uint32, C.c_float]
        L.or_outlier.argtypes = [P(or_view), C.c_int32, P(C.c_uint32), C.c_uint32, P(or_outlier_params),
                                 C.c_void_p, C.c_void_p, P(C.c_uint64), C.c_void_p, C.c_uint32,
                                 C.c_uint32, C.c_int]
        L.or_simple.argtypes = [P(or_view), C.c_int32, P(C.c_float), P(or_fade), C.c_int32, C.c_void_p,
                                C.c_uint32, C.c_uint32, C.c_int]
        L.or_shake.argtypes = [P(or_view), P(C.c_int32), C.c_uint32, C.c_uint32, C.c_uint32,
                               P(C.c_int32), C.c_int]
        L.or_crop_create.argtypes = [P(C.c_int32), C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_int32),
                                     P(C.c_uint32), P(C.c_uint32)]
        L.or_video_windows.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                       C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                       C.c_uint32, P(C.c_int32), P(C.c_int32), P(C.c_int32), C.c_int32]
        L.or_video_windows.restype = C.c_int32
        L.or_rng_range.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.or_rng_range.restype = C.c_uint32
        _lib = L
    return _lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def threshold(absolute: bool, lo: float, hi: float, bps: int = 1) -> or_threshold:
    t = or_threshold()
    lib().or_threshold_new(int(absolute), lo, hi, 255.0 if bps == 1 else 65535.0, C.byref(t))
    return t


def fade_table(frames, values):
    fr = np.ascontiguousarray(frames, np.int32)
    va = np.ascontiguousarray(values, np.float32)
    n = lib().or_fade_table(_p(fr, C.c_int32), _p(va, C.c_float), len(fr), None, 0)
    out = np.zeros(n, np.float32)
    lib().or_fade_table(_p(fr, C.c_int32), _p(va, C.c_float), len(fr), _p(out, C.c_float), n)
    return out


class Fade:
    """Oracle-side fade: keeps the table alive."""

    def __init__(self, repeat=False, absolute=True, frames=None, values=None, none=False):
        self.table = np.zeros(1, np.float32) if none else fade_table(frames, values)
        self.c = or_fade(int(none), int(repeat), int(absolute), 0 if none else int(frames[0]),
                         0 if none else len(self.table), _p(self.table, C.c_float))

    @classmethod
    def none(cls):
        return cls(none=True)

    def get(self, frame):
        return lib().or_fade_get(C.byref(self.c), frame)


class View:
    def __init__(self, frames: np.ndarray, slots=None, crop_xy=None, out_w=None, out_h=None):
        self.frames = np.ascontiguousarray(frames)
        nslot, h, w, c = self.frames.shape
        self.slots = None if slots is None else np.ascontiguousarray(slots, np.uint32)
        n = nslot if slots is None else len(self.slots)
        self.crop = None if crop_xy is None else np.ascontiguousarray(np.asarray(crop_xy, np.int32).reshape(-1))
        self.c = or_view(self.frames.ctypes.data, self.frames.dtype.itemsize, c, w, h,
                         _p(self.slots, C.c_uint32), n, _p(self.crop, C.c_int32),
                         out_w or w, out_h or h)
        self.out_shape = (out_h or h, out_w or w, c)


def outlier(frames, thr: or_threshold, background: int, outlier_mode: int, weights=(1, 1, 1, 1),
            fade: Fade = None, seed: int = 0, slots=None, crop_xy=None, out_w=None, out_h=None,
            frame_offset=None, sample_pos=None, threads=0, y0=0, y1=None, stats=False):
    """Returns (image, mask, warnings[, stats])."""
    v = View(frames, slots, crop_xy, out_w, out_h)
    fade = fade or Fade.none()
    p = or_outlier_params()
    p.threshold = thr
    p.background = background
    p.outlier = outlier_mode
    w = list(weights) + [1.0] * (4 - len(weights))
    for i in range(4):
        p.weights[i] = w[i]
    p.fade = fade.c
    p.seed = seed
    h, wd, c = v.out_shape
    y1 = h if y1 is None else y1
    img = np.zeros((h, wd, c), v.frames.dtype)
    mask = np.zeros((h, wd, c), np.uint8)
    st = np.zeros((h, wd), STATS_DTYPE) if stats else None
    warn = C.c_uint64(0)
    sp = None if sample_pos is None else np.ascontiguousarray(sample_pos, np.uint32)
    if frame_offset is None:
        frame_offset = int(slots[0]) if slots is not None else 0
    rc = lib().or_outlier(C.byref(v.c), frame_offset, _p(sp, C.c_uint32), 0 if sp is None else len(sp),
                          C.byref(p), img.ctypes.data, mask.ctypes.data, C.byref(warn),
                          None if st is None else st.ctypes.data, y0, y1, threads)
    if rc != 0:
        raise ValueError(f"oracle or_outlier returned {rc}")
    if stats:
        return img, mask, warn.value, st
    return img, mask, warn.value


def simple(frames, weights=(1, 1, 1, 1), fade: Fade = None, darker=True, slots=None, crop_xy=None,
           out_w=None, out_h=None, frame_offset=None, threads=0, y0=0, y1=None):
    v = View(frames, slots, crop_xy, out_w, out_h)
    fade = fade or Fade.none()
    w = np.array(list(weights) + [1.0] * (4 - len(weights)), np.float32)
    h, wd, c = v.out_shape
    y1 = h if y1 is None else y1
    img = np.zeros((h, wd, c), v.frames.dtype)
    if frame_offset is None:
        frame_offset = int(slots[0]) if slots is not None else 0
    rc = lib().or_simple(C.byref(v.c), frame_offset, _p(w, C.c_float), C.byref(fade.c), int(darker),
                         img.ctypes.data, y0, y1, threads)
    if rc != 0:
        raise ValueError(f"oracle or_simple returned {rc}")
    return img


def shake(frames, anchors, r, s, slots=None, threads=0):
    v = View(frames, slots)
    a = np.ascontiguousarray(np.asarray(anchors, np.int32).reshape(-1))
    off = np.zeros((v.c.n, 2), np.int32)
    rc = lib().or_shake(C.byref(v.c), _p(a, C.c_int32), len(a) // 2, r, s, _p(off, C.c_int32), threads)
    if rc != 0:
        raise ValueError(f"oracle or_shake returned {rc}")
    return off


def crop_create(offsets, w, h):
    off = np.ascontiguousarray(np.asarray(offsets, np.int32).reshape(-1, 2))
    crop = np.zeros_like(off)
    ow, oh = C.c_uint32(), C.c_uint32()
    has = lib().or_crop_create(_p(off, C.c_int32), len(off), w, h, _p(crop, C.c_int32), C.byref(ow),
                               C.byref(oh))
    return (crop, ow.value, oh.value) if has else None


def video_windows(vin, vout, image_count):
    """vin/vout: (start|None, end|None, step)."""
    args = []
    for r in (vin, vout):
        s, e, st = r
        args += [int(s is not None), s or 0, int(e is not None), e or 0, st]
    n = lib().or_video_windows(*args, image_count, None, None, None, 0)
    cnt = max(n, 1)
    a, b, f = (np.zeros(cnt, np.int32) for _ in range(3))
    lib().or_video_windows(*args, image_count, _p(a, C.c_int32), _p(b, C.c_int32), _p(f, C.c_int32), cnt)
    return [(int(f[i]), list(range(int(a[i]), int(b[i]), vin[2]))) for i in range(n)]


def median(sorted_vals):
    a = np.ascontiguousarray(sorted_vals, np.uint32)
    return lib().or_median(_p(a, C.c_uint32), len(a))


def quantile(sorted_vals, q):
    a = np.ascontiguousarray(sorted_vals, np.uint32)
    return lib().or_quantile(_p(a, C.c_uint32), len(a), q)


def blend_into_u8(a, b, blend):
    a = np.ascontiguousarray(a, np.uint8).copy()
    b = np.ascontiguousarray(b, np.uint8)
    lib().or_blend_into_u8(_p(a, C.c_uint8), _p(b, C.c_uint8), len(a), blend)
    return a
