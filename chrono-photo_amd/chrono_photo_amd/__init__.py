"""chrono_photo_amd -- Python binding of the MI355X compositing library (ctypes).

Mirrors the reference crate's library API for the compositing path (chrono-photo v0.6.5):

  Threshold            src/options.rs:189-281   (abs|rel)/min[/max]
  Fade                 src/options.rs:60-186    (clamp|repeat)/(abs|rel)/f,v/f,v...
  BackgroundMode       src/options.rs:319-345
  OutlierSelectionMode src/options.rs:284-316
  SelectionMode        src/options.rs:8-32
  OutlierProcessor     src/chrono.rs:32-206     new(...).process(...) -> (image, blend mask)
  SimpleProcessor      src/simple.rs:11-168     new(weights, fade, darker).process(...)
  ShakeAnalyzer        src/shake.rs:184-313     analyze(...) -> offsets
  Crop                 src/shake.rs:121-181     Crop.create(offsets, width, height)
  FrameRange           src/flist.rs:9-79        start/end/step with "." placeholders
  video_windows        src/main.rs:230-286

Frames live in a device-resident Stack instead of image files and gzip slice files.  All
compute runs in libchrono_hip.so (HIP, gfx950); importing this module fails loudly when
that library has not been built, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")


class ChronoError(RuntimeError):
    """A non-zero cp_status from the library (message from cp_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


CP_OK = 0
CP_ERR_INVALID_ARG = -1
CP_ERR_LAYOUT_MISMATCH = -2
CP_ERR_ANCHOR_OUT_OF_RANGE = -3
CP_ERR_TOO_FEW_FRAMES_REL = -4
CP_ERR_OUT_OF_HBM_BUDGET = -5
CP_ERR_HIP = -6
CP_ERR_NO_DEVICE = -7
CP_ERR_CAPACITY = -8

CP_FLAG_DEVICE_OUT = 1
CP_FLAG_FORCE_GENERIC = 2
CP_FLAG_NO_GRAPH = 4

CP_MODE_OUTLIER = 0
CP_MODE_LIGHTER = 1
CP_MODE_DARKER = 2


class cp_threshold(C.Structure):
    _fields_ = [("absolute", C.c_int32), ("min", C.c_float), ("max", C.c_float), ("scale", C.c_float)]


class cp_fade(C.Structure):
    _fields_ = [("is_none", C.c_int32), ("repeat", C.c_int32), ("absolute", C.c_int32),
                ("offset", C.c_int32), ("len", C.c_uint32), ("values", C.POINTER(C.c_float))]


class cp_outlier_params(C.Structure):
    _fields_ = [("threshold", cp_threshold), ("background", C.c_int32), ("outlier", C.c_int32),
                ("weights", C.c_float * 4), ("fade", cp_fade), ("seed", C.c_uint64),
                ("sample_pos", C.POINTER(C.c_uint32)), ("n_sample", C.c_uint32)]


class cp_frames(C.Structure):
    _fields_ = [("frames", C.POINTER(C.c_uint32)), ("n", C.c_uint32), ("frame_offset", C.c_int32),
                ("crop_xy", C.POINTER(C.c_int32)), ("out_width", C.c_uint32),
                ("out_height", C.c_uint32)]


class cp_pixel_stats(C.Structure):
    _fields_ = [("median", C.c_float * 4), ("iqr_inv", C.c_float * 4), ("num_outliers", C.c_uint32),
                ("max_index", C.c_uint32), ("max_dist_sq", C.c_float), ("bg_index", C.c_uint32),
                ("warning", C.c_uint32), ("pad", C.c_uint32)]


PIXEL_STATS_DTYPE = np.dtype([("median", "<f4", (4,)), ("iqr_inv", "<f4", (4,)),
                              ("num_outliers", "<u4"), ("max_index", "<u4"), ("max_dist_sq", "<f4"),
                              ("bg_index", "<u4"), ("warning", "<u4"), ("pad", "<u4")])


class cp_outputs(C.Structure):
    _fields_ = [("px", C.c_void_p), ("mask", C.c_void_p), ("warnings", C.c_void_p),
                ("stats", C.c_void_p)]


class cp_stack_info(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("bytes_per_sample", C.c_uint32), ("capacity", C.c_uint32), ("count", C.c_uint32),
                ("device_ptr", C.c_void_p)]


class cp_frame_range(C.Structure):
    _fields_ = [("has_start", C.c_int32), ("start", C.c_int32), ("has_end", C.c_int32),
                ("end", C.c_int32), ("step", C.c_uint32)]


class cp_job(C.Structure):
    _fields_ = [("mode", C.c_int32), ("stack", C.c_void_p), ("frames", cp_frames),
                ("outlier", cp_outlier_params), ("weights", C.c_float * 4), ("fade", cp_fade),
                ("out", cp_outputs)]


class cp_request(C.Structure):
    _fields_ = [("job", cp_job), ("host_frames", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("channels", C.c_uint32), ("bytes_per_sample", C.c_uint32), ("n_frames", C.c_uint32),
                ("host_out_px", C.c_void_p), ("host_out_mask", C.c_void_p)]


class cp_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("pixel_samples", C.c_uint64), ("peak_bytes", C.c_uint64), ("peak_inflight", C.c_uint32),
                ("kernel_downloads", C.c_uint32)]


class cp_batch_info(C.Structure):
    _fields_ = [("group_launches", C.c_uint32), ("group_jobs", C.c_uint32),
                ("direct_jobs", C.c_uint32), ("reserved", C.c_uint32)]


class cs_scene(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("channels", C.c_uint32), ("bytes_per_sample", C.c_uint32), ("n_frames", C.c_uint32),
                ("base_lo", C.c_int32), ("base_hi", C.c_int32), ("sigma_x16", C.c_int32),
                ("n_disks", C.c_uint32), ("disk_r_min", C.c_int32), ("disk_r_max", C.c_int32),
                ("disk_val_lo", C.c_int32), ("disk_val_hi", C.c_int32), ("blob_permille", C.c_uint32),
                ("blob_r", C.c_int32), ("blob_val", C.c_int32), ("shake_max", C.c_int32),
                ("texture_amp", C.c_int32), ("disk_y_lo", C.c_int32), ("disk_y_hi", C.c_int32)]


# Every symbol include/chrono_hip.h declares (checked by tests/test_abi.py).
ABI_SYMBOLS = [
    "cp_last_error", "cp_version", "cp_ctx_create", "cp_ctx_destroy", "cp_ctx_synchronize",
    "cp_ctx_stream", "cp_stack_create", "cp_stack_wrap", "cp_stack_upload", "cp_stack_push",
    "cp_stack_commit", "cp_stack_download",
    "cp_stack_info_get", "cp_stack_destroy", "cp_threshold_new", "cp_threshold_blend_value",
    "cp_fade_table", "cp_fade_get", "cp_outlier", "cp_outlier_host", "cp_simple", "cp_simple_host", "cp_shake",
    "cp_shake_noalloc", "cp_crop_create",
    "cp_video_windows", "cp_batch_run", "cp_batch_create", "cp_batch_launch", "cp_batch_destroy",
    "cp_batch_info_get", "cp_autotune", "cp_tile_shapes_set", "cp_server_create", "cp_server_submit", "cp_server_poll",
    "cp_server_wait", "cp_server_stats_get", "cp_server_destroy",
]

_lib = None
_synth = None


def lib_path(name: str = "libchrono_hip.so") -> str:
    return os.path.join(LIB_DIR, name)


def load_library() -> C.CDLL:
    """Loads libchrono_hip.so; raises when it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python chrono-photo_amd/build.py` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "cp_last_error": (C.c_char_p, []),
        "cp_version": (C.c_char_p, []),
        "cp_ctx_create": (C.c_int, [C.c_int32, C.c_uint64, P(C.c_void_p)]),
        "cp_ctx_destroy": (C.c_int, [C.c_void_p]),
        "cp_ctx_synchronize": (C.c_int, [C.c_void_p]),
        "cp_ctx_stream": (C.c_void_p, [C.c_void_p]),
        "cp_stack_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, P(C.c_void_p)]),
        "cp_stack_wrap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, P(C.c_void_p)]),
        "cp_stack_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
        "cp_stack_push": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "cp_stack_commit": (C.c_int, [C.c_void_p, C.c_uint32]),
        "cp_stack_download": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
        "cp_stack_info_get": (C.c_int, [C.c_void_p, P(cp_stack_info)]),
        "cp_stack_destroy": (C.c_int, [C.c_void_p]),
        "cp_threshold_new": (C.c_int, [C.c_int32, C.c_float, C.c_float, C.c_uint32, P(cp_threshold)]),
        "cp_threshold_blend_value": (C.c_float, [P(cp_threshold), C.c_float]),
        "cp_fade_table": (C.c_int, [P(C.c_int32), P(C.c_float), C.c_uint32, P(C.c_float), C.c_uint32,
                                    P(C.c_uint32)]),
        "cp_fade_get": (C.c_float, [P(cp_fade), C.c_int32]),
        "cp_outlier": (C.c_int, [C.c_void_p, C.c_void_p, P(cp_frames), P(cp_outlier_params),
                                 P(cp_outputs), C.c_uint32]),
        "cp_outlier_host": (C.c_int, [C.c_void_p, P(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, P(cp_frames), P(cp_outlier_params), P(cp_outputs), C.c_uint64]),
        "cp_simple": (C.c_int, [C.c_void_p, C.c_void_p, P(cp_frames), P(C.c_float), P(cp_fade),
                                C.c_int32, C.c_void_p, C.c_uint32]),
        "cp_simple_host": (C.c_int, [C.c_void_p, P(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, P(cp_frames), P(C.c_float), P(cp_fade), C.c_int32, C.c_void_p,
                                     C.c_uint64]),
        "cp_shake": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_uint32), C.c_uint32, P(C.c_int32),
                               C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_int32)]),
        "cp_shake_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_uint32), C.c_uint32, P(C.c_int32),
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]),
        "cp_crop_create": (C.c_int, [P(C.c_int32), C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_int32),
                                     P(C.c_uint32), P(C.c_uint32), P(C.c_int32)]),
        "cp_video_windows": (C.c_int, [P(cp_frame_range), P(cp_frame_range), C.c_uint32, P(C.c_int32),
                                       P(C.c_int32), P(C.c_int32), C.c_int32, P(C.c_int32)]),
        "cp_batch_run": (C.c_int, [C.c_void_p, P(cp_job), C.c_uint32, C.c_uint32]),
        "cp_batch_create": (C.c_int, [C.c_void_p, P(cp_job), C.c_uint32, C.c_uint32, P(C.c_void_p)]),
        "cp_batch_launch": (C.c_int, [C.c_void_p]),
        "cp_batch_destroy": (C.c_int, [C.c_void_p]),
        "cp_batch_info_get": (C.c_int, [C.c_void_p, P(cp_batch_info)]),
        "cp_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_int32)]),
        "cp_tile_shapes_set": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_int32)]),
        "cp_server_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, P(C.c_void_p)]),
        "cp_server_submit": (C.c_int, [C.c_void_p, P(cp_request), P(C.c_uint64)]),
        "cp_server_stats_get": (C.c_int, [C.c_void_p, P(cp_server_stats)]),
        "cp_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, P(C.c_int32)]),
        "cp_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, P(C.c_double)]),
        "cp_server_destroy": (C.c_int, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def load_synth() -> C.CDLL:
    global _synth
    if _synth is not None:
        return _synth
    path = lib_path("libchrono_synth.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python chrono-photo_amd/build.py`")
    s = C.CDLL(path)
    s.cs_generate_host.restype = C.c_int
    s.cs_generate_host.argtypes = [C.POINTER(cs_scene), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_void_p]
    s.cs_generate_device.restype = C.c_int
    s.cs_generate_device.argtypes = [C.POINTER(cs_scene), C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    s.cs_shake_truth.restype = None
    s.cs_shake_truth.argtypes = [C.POINTER(cs_scene), C.c_uint32, C.POINTER(C.c_int32)]
    _synth = s
    return s


def _check(rc: int):
    if rc != CP_OK:
        msg = load_library().cp_last_error()
        raise ChronoError(rc, msg.decode() if msg else "")


def _ptr(a: Optional[np.ndarray], ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


# ----------------------------------------------------------------- option values

class SelectionMode(enum.IntEnum):
    """src/options.rs:8-32"""
    OUTLIER = 0
    LIGHTER = 1
    DARKER = 2

    @classmethod
    def from_str(cls, s: str) -> "SelectionMode":
        table = {"lighter": cls.LIGHTER, "darker": cls.DARKER, "outlier": cls.OUTLIER}
        if s not in table:
            raise ValueError(f"Not a pixel selection mode: {s}. Must be one of (lighter|darker|outlier)")
        return table[s]


class BackgroundMode(enum.IntEnum):
    """src/options.rs:319-345"""
    FIRST = 0
    RANDOM = 1
    AVERAGE = 2
    MEDIAN = 3

    @classmethod
    def from_str(cls, s: str) -> "BackgroundMode":
        table = {"first": cls.FIRST, "random": cls.RANDOM, "average": cls.AVERAGE, "median": cls.MEDIAN}
        if s not in table:
            raise ValueError(f"Not a background pixel selection mode: {s}. "
                             "Must be one of (first|random|average|median)")
        return table[s]


class OutlierSelectionMode(enum.IntEnum):
    """src/options.rs:284-316"""
    FIRST = 0
    LAST = 1
    EXTREME = 2
    AVERAGE = 3
    FORWARD = 4
    BACKWARD = 5

    @classmethod
    def from_str(cls, s: str) -> "OutlierSelectionMode":
        table = {"first": cls.FIRST, "last": cls.LAST, "extreme": cls.EXTREME, "average": cls.AVERAGE,
                 "forward": cls.FORWARD, "backward": cls.BACKWARD}
        if s not in table:
            raise ValueError(f"Not an outlier selection mode: {s}. "
                             "Must be one of (first|last|extreme|average|forward|backward)")
        return table[s]


@dataclass
class Threshold:
    """Threshold (src/options.rs:189-281).  `absolute`, `min`, `max` are the parsed values;
    the pre-scaled values of Threshold::new come from `native(bytes_per_sample)`."""
    absolute: bool
    min: float
    max: float

    @classmethod
    def abs(cls, min: float, max: float) -> "Threshold":
        return cls(True, min, max)

    @classmethod
    def rel(cls, min: float, max: float) -> "Threshold":
        return cls(False, min, max)

    @classmethod
    def from_str(cls, s: str) -> "Threshold":
        parts = s.split("/")
        if parts[0] in ("absolute", "abs"):
            absolute = True
        elif parts[0] in ("relative", "rel"):
            absolute = False
        else:
            raise ValueError(f"Not a pixel outlier detection mode: {s}. "
                             "Must be one of (abs[olute]|rel[ative])/<min>[/<max>]")
        if len(parts) < 2:
            raise ValueError(f"Unexpected format in {s}")
        lo = float(np.float32(parts[1]))
        hi = float(np.float32(parts[2])) if len(parts) > 2 else lo
        return cls(absolute, lo, hi)

    def native(self, bytes_per_sample: int = 1) -> cp_threshold:
        t = cp_threshold()
        _check(load_library().cp_threshold_new(int(self.absolute), self.min, self.max,
                                               bytes_per_sample, C.byref(t)))
        return t

    def blend_value(self, dist: float, bytes_per_sample: int = 1) -> float:
        t = self.native(bytes_per_sample)
        return load_library().cp_threshold_blend_value(C.byref(t), dist)


class Fade:
    """Fade (src/options.rs:60-186)."""

    def __init__(self, repeat: bool, absolute: bool, frames: Sequence[tuple], _none: bool = False):
        self.is_none = _none
        self.repeat = repeat
        self.absolute = absolute
        if _none:
            self.offset = 0
            self.values = np.zeros(0, np.float32)
            return
        if len(frames) < 2:
            raise ValueError("Fade requires at least two frames specified.")
        fr = np.array([f for f, _ in frames], np.int32)
        vals = np.array([v for _, v in frames], np.float32)
        n = C.c_uint32(0)
        _check(load_library().cp_fade_table(_ptr(fr, C.c_int32), _ptr(vals, C.c_float), len(fr),
                                            None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(load_library().cp_fade_table(_ptr(fr, C.c_int32), _ptr(vals, C.c_float), len(fr),
                                            _ptr(out, C.c_float), n.value, C.byref(n)))
        self.offset = int(fr[0])
        self.values = out

    @classmethod
    def new(cls, mode: str, absolute: bool, frames) -> "Fade":
        return cls(mode == "repeat", absolute, frames)

    @classmethod
    def none(cls) -> "Fade":
        return cls(False, True, [], _none=True)

    @classmethod
    def from_str(cls, s: str) -> "Fade":
        parts = s.split("/")
        if parts[0] not in ("repeat", "clamp"):
            raise ValueError(f"Not a fade mode: {parts[0]}. Must be one of (repeat|clamp)")
        if len(parts) < 2 or parts[1] not in ("absolute", "abs", "relative", "rel"):
            raise ValueError(f"Not a frame fade spec: {s}")
        frames = []
        for p in parts[2:]:
            fv = p.split(",")
            if len(fv) != 2:
                raise ValueError(f"Expected (int,float) per frame for fade. Got: {s}")
            frames.append((int(fv[0]), float(np.float32(fv[1]))))
        return cls(parts[0] == "repeat", parts[1] in ("absolute", "abs"), frames)

    def native(self) -> cp_fade:
        f = cp_fade()
        f.is_none = int(self.is_none)
        f.repeat = int(self.repeat)
        f.absolute = int(self.absolute)
        f.offset = self.offset
        f.len = len(self.values)
        f.values = _ptr(self.values, C.c_float) if len(self.values) else None
        return f

    def get(self, frame: int) -> float:
        f = self.native()
        return load_library().cp_fade_get(C.byref(f), frame)


@dataclass
class FrameRange:
    """FrameRange (src/flist.rs:9-79)."""
    start: Optional[int] = None
    end: Optional[int] = None
    step: int = 1

    @classmethod
    def from_str(cls, s: str) -> "FrameRange":
        parts = s.split("/")
        if len(parts) != 3:
            raise ValueError(f"Option --frames expects 3 elements: start/end/step, {len(parts)} were suppied")
        vals = [None if p == "." else int(p) for p in parts]
        return cls(vals[0], vals[1], vals[2] if vals[2] is not None else 1)

    def native(self) -> cp_frame_range:
        r = cp_frame_range()
        r.has_start = int(self.start is not None)
        r.start = self.start or 0
        r.has_end = int(self.end is not None)
        r.end = self.end or 0
        r.step = self.step
        return r


def video_windows(video_in: FrameRange, video_out: FrameRange, image_count: int):
    """Window table of create_video (src/main.rs:230-286): list of (frame, [indices])."""
    lib = load_library()
    n = C.c_int32(0)
    vi, vo = video_in.native(), video_out.native()
    _check(lib.cp_video_windows(C.byref(vi), C.byref(vo), image_count, None, None, None, 0, C.byref(n)))
    cnt = max(n.value, 1)
    st, en, fr = (np.zeros(cnt, np.int32) for _ in range(3))
    _check(lib.cp_video_windows(C.byref(vi), C.byref(vo), image_count, _ptr(st, C.c_int32),
                                _ptr(en, C.c_int32), _ptr(fr, C.c_int32), cnt, C.byref(n)))
    out = []
    for i in range(n.value):
        out.append((int(fr[i]), list(range(int(st[i]), int(en[i]), video_in.step))))
    return out


# ------------------------------------------------------------------- runtime

class Context:
    """A device context (stream, HBM budget).  Not thread-safe."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        self.handle = None
        self._children = weakref.WeakSet()
        _check(self.lib.cp_ctx_create(device, hbm_budget, C.byref(h)))
        self.handle = h

    def synchronize(self):
        _check(self.lib.cp_ctx_synchronize(self.handle))

    @property
    def stream(self) -> int:
        return self.lib.cp_ctx_stream(self.handle)

    def close(self):
        """Destroys the context; stacks and servers created on it are closed first."""
        if self.handle:
            for child in list(self._children):
                child.close()
            self.lib.cp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stack:
    """Device-resident frame stack [capacity][height][width][channels] (u8 or u16)."""

    def __init__(self, ctx: Context, width: int, height: int, channels: int, bytes_per_sample: int = 1,
                 capacity: int = 1, device_ptr: Optional[int] = None, n_frames: Optional[int] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = None
        h = C.c_void_p()
        if device_ptr is None:
            _check(self.lib.cp_stack_create(ctx.handle, width, height, channels, bytes_per_sample,
                                            capacity, C.byref(h)))
        else:
            _check(self.lib.cp_stack_wrap(ctx.handle, C.c_void_p(device_ptr), width, height, channels,
                                          bytes_per_sample, n_frames, C.byref(h)))
        self.handle = h
        ctx._children.add(self)
        self.width, self.height, self.channels, self.bps = width, height, channels, bytes_per_sample

    @classmethod
    def from_array(cls, ctx: Context, frames: np.ndarray, capacity: Optional[int] = None) -> "Stack":
        """frames: [n][h][w][c] uint8 or uint16."""
        frames = np.ascontiguousarray(frames)
        if frames.ndim != 4:
            raise ValueError("frames must be [n][h][w][c]")
        n, h, w, c = frames.shape
        bps = frames.dtype.itemsize
        st = cls(ctx, w, h, c, bps, capacity or n)
        st.upload(0, frames)
        return st

    @property
    def dtype(self):
        return np.uint8 if self.bps == 1 else np.uint16

    def upload(self, slot: int, frames: np.ndarray):
        frames = np.ascontiguousarray(frames, dtype=self.dtype)
        _check(self.lib.cp_stack_upload(self.handle, slot, frames.ctypes.data_as(C.c_void_p),
                                        frames.shape[0]))

    def push(self, frames: np.ndarray) -> int:
        frames = np.ascontiguousarray(frames, dtype=self.dtype)
        first = C.c_uint32(0)
        _check(self.lib.cp_stack_push(self.handle, frames.ctypes.data_as(C.c_void_p), frames.shape[0],
                                      C.byref(first)))
        return first.value

    def download(self, slot: int = 0, n: Optional[int] = None) -> np.ndarray:
        """Copies frames [slot, slot + n) back to the host."""
        n = self.count - slot if n is None else n
        out = np.zeros((n, self.height, self.width, self.channels), self.dtype)
        _check(self.lib.cp_stack_download(self.handle, slot, out.ctypes.data_as(C.c_void_p), n))
        return out

    def commit(self, count: int):
        """Declares slots [0, count) filled by a device-side writer."""
        _check(self.lib.cp_stack_commit(self.handle, count))

    def info(self) -> cp_stack_info:
        i = cp_stack_info()
        _check(self.lib.cp_stack_info_get(self.handle, C.byref(i)))
        return i

    @property
    def count(self) -> int:
        return self.info().count

    @property
    def device_ptr(self) -> int:
        return self.info().device_ptr

    def close(self):
        if self.handle and self.ctx.handle:
            self.lib.cp_stack_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _FrameSel:
    """Keeps the arrays behind a cp_frames alive."""

    def __init__(self, stack: Optional[Stack], image_indices=None, crop=None, n=None):
        if image_indices is None:
            n = stack.count if n is None else n
            self.idx = None
            self.offset = 0
        else:
            self.idx = np.ascontiguousarray(image_indices, np.uint32)
            n = len(self.idx)
            self.offset = int(self.idx[0]) if n else 0
        self.n = n
        self.crop_xy = None
        self.out_w = self.out_h = 0
        if crop is not None:
            crop_xy, ow, oh = crop
            self.crop_xy = np.ascontiguousarray(crop_xy, np.int32).reshape(-1)
            self.out_w, self.out_h = ow, oh
        self.c = cp_frames()
        self.c.frames = _ptr(self.idx, C.c_uint32)
        self.c.n = n
        self.c.frame_offset = self.offset
        self.c.crop_xy = _ptr(self.crop_xy, C.c_int32)
        self.c.out_width = self.out_w
        self.c.out_height = self.out_h

    def out_shape(self, stack: Stack):
        w = self.out_w or stack.width
        h = self.out_h or stack.height
        return h, w, stack.channels


def _host_frames(frames):
    """Contiguous host frames of one [h][w][c] u8/u16 layout and a pointer array to them."""
    arrs = [np.ascontiguousarray(f) for f in frames]
    if not arrs:
        raise ValueError("no frames")
    shape, dtype = arrs[0].shape, arrs[0].dtype
    if len(shape) != 3 or dtype not in (np.uint8, np.uint16) or any(a.shape != shape or a.dtype != dtype
                                                                      for a in arrs):
        raise ValueError("frames must share one [h][w][c] u8 or u16 layout")
    return arrs, (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs]), shape, dtype


class OutlierProcessor:
    """OutlierProcessor (src/chrono.rs:32-206) over a device-resident stack."""

    def __init__(self, threshold: Threshold, bg_mode: BackgroundMode, outlier_mode: OutlierSelectionMode,
                 weights=(1.0, 1.0, 1.0, 1.0), fade: Optional[Fade] = None,
                 sample_indices: Optional[Sequence[int]] = None, seed: int = 0):
        self.threshold = threshold
        self.background = BackgroundMode(bg_mode)
        self.outlier = OutlierSelectionMode(outlier_mode)
        self.weights = np.array(list(weights) + [1.0] * (4 - len(weights)), np.float32)
        self.fade = fade or Fade.none()
        self.sample_indices = None if sample_indices is None else np.ascontiguousarray(sample_indices, np.uint32)
        self.seed = seed

    def params(self, bytes_per_sample: int) -> cp_outlier_params:
        p = cp_outlier_params()
        p.threshold = self.threshold.native(bytes_per_sample)
        p.background = int(self.background)
        p.outlier = int(self.outlier)
        for i in range(4):
            p.weights[i] = float(self.weights[i])
        p.fade = self.fade.native()
        p.seed = self.seed
        p.sample_pos = _ptr(self.sample_indices, C.c_uint32)
        p.n_sample = 0 if self.sample_indices is None else len(self.sample_indices)
        return p

    def process(self, stack: Stack, image_indices=None, crop=None, with_mask: bool = True,
                with_stats: bool = False, force_generic: bool = False):
        """Returns (image, blend_mask[, stats]) as numpy arrays [h][w][c]; the warning count
        is kept in `self.warnings`."""
        sel = _FrameSel(stack, image_indices, crop)
        h, w, c = sel.out_shape(stack)
        img = np.zeros((h, w, c), stack.dtype)
        mask = np.zeros((h, w, c), np.uint8) if with_mask else None
        stats = np.zeros((h, w), PIXEL_STATS_DTYPE) if with_stats else None
        warn = np.zeros(1, np.uint64)
        out = cp_outputs(img.ctypes.data, mask.ctypes.data if mask is not None else None,
                         warn.ctypes.data, stats.ctypes.data if stats is not None else None)
        p = self.params(stack.bps)
        flags = CP_FLAG_FORCE_GENERIC if force_generic else 0
        _check(stack.lib.cp_outlier(stack.ctx.handle, stack.handle, C.byref(sel.c), C.byref(p),
                                    C.byref(out), flags))
        self.warnings = int(warn[0])
        res = (img, mask)
        if with_stats:
            res = res + (stats,)
        return res

    def process_host(self, ctx: "Context", frames, image_indices=None, crop=None, with_mask: bool = True,
                     with_stats: bool = False, band_bytes: int = 0):
        """Out-of-core composite of frames held in host memory (cp_outlier_host): `frames` is
        an array [n][h][w][c] or a sequence of [h][w][c] arrays (u8 or u16); row bands stream
        through a device band stack of at most `band_bytes` (0: what the context allows).
        Returns the same as process() over a stack of these frames."""
        arrs, ptrs, (h0, w0, c0), dtype = _host_frames(frames)
        n = len(arrs) if image_indices is None else None
        sel = _FrameSel(None, image_indices, crop, n=n)
        h = sel.out_h or h0
        w = sel.out_w or w0
        img = np.zeros((h, w, c0), dtype)
        mask = np.zeros((h, w, c0), np.uint8) if with_mask else None
        stats = np.zeros((h, w), PIXEL_STATS_DTYPE) if with_stats else None
        warn = np.zeros(1, np.uint64)
        out = cp_outputs(img.ctypes.data, mask.ctypes.data if mask is not None else None,
                         warn.ctypes.data, stats.ctypes.data if stats is not None else None)
        bps = dtype.itemsize
        p = self.params(bps)
        _check(ctx.lib.cp_outlier_host(ctx.handle, ptrs, len(arrs), w0, h0, c0, bps, C.byref(sel.c),
                                       C.byref(p), C.byref(out), band_bytes))
        self.warnings = int(warn[0])
        res = (img, mask)
        if with_stats:
            res = res + (stats,)
        return res


class SimpleProcessor:
    """SimpleProcessor (src/simple.rs:11-168)."""

    def __init__(self, weights=(1.0, 1.0, 1.0, 1.0), fade: Optional[Fade] = None, darker: bool = True):
        self.weights = np.array(list(weights) + [1.0] * (4 - len(weights)), np.float32)
        self.fade = fade or Fade.none()
        self.darker = darker

    def process(self, stack: Stack, crop=None, image_indices=None):
        sel = _FrameSel(stack, image_indices, crop)
        h, w, c = sel.out_shape(stack)
        img = np.zeros((h, w, c), stack.dtype)
        f = self.fade.native()
        _check(stack.lib.cp_simple(stack.ctx.handle, stack.handle, C.byref(sel.c),
                                   _ptr(self.weights, C.c_float), C.byref(f), int(self.darker),
                                   C.c_void_p(img.ctypes.data), 0))
        return img

    def process_host(self, ctx: "Context", frames, crop=None, image_indices=None, band_bytes: int = 0):
        """Out-of-core composite of frames held in host memory (cp_simple_host); returns the
        same image as process() over a stack of these frames."""
        arrs, ptrs, (h0, w0, c0), dtype = _host_frames(frames)
        sel = _FrameSel(None, image_indices, crop, n=len(arrs) if image_indices is None else None)
        img = np.zeros((sel.out_h or h0, sel.out_w or w0, c0), dtype)
        f = self.fade.native()
        _check(ctx.lib.cp_simple_host(ctx.handle, ptrs, len(arrs), w0, h0, c0, dtype.itemsize, C.byref(sel.c),
                                      _ptr(self.weights, C.c_float), C.byref(f), int(self.darker),
                                      C.c_void_p(img.ctypes.data), band_bytes))
        return img


class ShakeAnalyzer:
    """ShakeAnalyzer (src/shake.rs:184-313)."""

    def analyze(self, stack: Stack, anchors, anchor_radius: int, search_radius: int, frames=None):
        n = stack.count if frames is None else len(frames)
        idx = None if frames is None else np.ascontiguousarray(frames, np.uint32)
        anc = np.ascontiguousarray(np.asarray(anchors, np.int32).reshape(-1))
        offsets = np.zeros((n, 2), np.int32)
        _check(stack.lib.cp_shake(stack.ctx.handle, stack.handle, _ptr(idx, C.c_uint32), n,
                                  _ptr(anc, C.c_int32), len(anc) // 2, anchor_radius, search_radius,
                                  _ptr(offsets, C.c_int32)))
        return offsets


class Crop:
    """Crop::create (src/shake.rs:136-176).  Returns None when every offset is zero,
    else (crop_xy [n][2], out_w, out_h)."""

    @staticmethod
    def create(offsets, width: int, height: int):
        off = np.ascontiguousarray(np.asarray(offsets, np.int32).reshape(-1, 2))
        crop = np.zeros_like(off)
        ow, oh, has = C.c_uint32(), C.c_uint32(), C.c_int32()
        _check(load_library().cp_crop_create(_ptr(off, C.c_int32), len(off), width, height,
                                             _ptr(crop, C.c_int32), C.byref(ow), C.byref(oh), C.byref(has)))
        if not has.value:
            return None
        return crop, ow.value, oh.value


# ------------------------------------------------------------ synthetic scenes

@dataclass
class Scene:
    """Seeded synthetic stack (csrc/synth.h)."""
    seed: int
    width: int
    height: int
    channels: int = 3
    bytes_per_sample: int = 1
    n_frames: int = 16
    base_lo: int = 40
    base_hi: int = 200
    sigma: float = 3.0
    n_disks: int = 1
    disk_r_min: int = 20
    disk_r_max: int = 20
    disk_val_lo: int = 15
    disk_val_hi: int = 25
    blob_permille: int = 0
    blob_r: int = 30
    blob_val: int = 230
    shake_max: int = 0
    texture_amp: int = 0
    disk_y_lo: int = 0
    disk_y_hi: int = 0

    def native(self) -> cs_scene:
        s = cs_scene()
        for f, _ in cs_scene._fields_:
            if f == "sigma_x16":
                s.sigma_x16 = int(round(self.sigma * 16))
            else:
                setattr(s, f, getattr(self, f))
        return s

    @property
    def dtype(self):
        return np.uint8 if self.bytes_per_sample == 1 else np.uint16

    def generate(self, first: int = 0, n: Optional[int] = None, y0: int = 0, y1: Optional[int] = None):
        n = self.n_frames - first if n is None else n
        y1 = self.height if y1 is None else y1
        out = np.zeros((n, y1 - y0, self.width, self.channels), self.dtype)
        s = self.native()
        rc = load_synth().cs_generate_host(C.byref(s), first, n, y0, y1, out.ctypes.data)
        if rc:
            raise ChronoError(rc, "cs_generate_host failed")
        return out

    def generate_device(self, dev_ptr: int, first: int = 0, n: Optional[int] = None, stream: int = 0):
        n = self.n_frames - first if n is None else n
        s = self.native()
        rc = load_synth().cs_generate_device(C.byref(s), first, n, C.c_void_p(dev_ptr), C.c_void_p(stream))
        if rc:
            raise ChronoError(rc, "cs_generate_device failed")

    def shake_truth(self, n: Optional[int] = None):
        n = self.n_frames if n is None else n
        out = np.zeros((n, 2), np.int32)
        load_synth().cs_shake_truth(C.byref(self.native()), n, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out


# ------------------------------------------------------------------- batching

class Job:
    """One composite for a batch or the server: a mode, a stack, a frame selection and
    device output pointers (out_px, out_mask, warnings: device addresses or None)."""

    def __init__(self, mode: int, stack: Optional[Stack], out_px: Optional[int], out_mask: Optional[int] = None,
                 warnings: Optional[int] = None, processor=None, image_indices=None, crop=None,
                 n_frames: Optional[int] = None, bytes_per_sample: Optional[int] = None):
        self.sel = _FrameSel(stack, image_indices, crop, n_frames)
        self.stack = stack
        self.processor = processor
        self.c = cp_job()
        self.c.mode = mode
        self.c.stack = stack.handle if stack is not None else None
        self.c.frames = self.sel.c
        bps = stack.bps if stack is not None else bytes_per_sample
        if mode == CP_MODE_OUTLIER:
            self.c.outlier = processor.params(bps)
        else:
            proc = processor or SimpleProcessor(darker=mode == CP_MODE_DARKER)
            self.processor = proc
            for i in range(4):
                self.c.weights[i] = float(proc.weights[i])
            self.c.fade = proc.fade.native()
        self.c.out = cp_outputs(out_px, out_mask, warnings, None)

    def out_shape(self):
        return self.sel.out_shape(self.stack)


def _job_array(jobs):
    arr = (cp_job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.c
    return arr


class Batch:
    """cp_batch: many composites, one launch per register-resident kernel shape the jobs are
    planned on (8-bit 1-4 channels up to 1280 frames, 16-bit 3-4 channels up to 512) covering
    all their tiles, the other jobs launched one by one, spread over four streams and captured
    in a HIP graph."""

    def __init__(self, ctx: Context, jobs, graph: bool = True):
        self.ctx = ctx
        self.lib = ctx.lib
        self.jobs = list(jobs)
        self.handle = None
        arr = _job_array(self.jobs)
        h = C.c_void_p()
        _check(self.lib.cp_batch_create(ctx.handle, arr, len(self.jobs), 0 if graph else CP_FLAG_NO_GRAPH,
                                        C.byref(h)))
        self.handle = h
        ctx._children.add(self)

    def launch(self):
        _check(self.lib.cp_batch_launch(self.handle))

    def info(self) -> dict:
        i = cp_batch_info()
        _check(self.lib.cp_batch_info_get(self.handle, C.byref(i)))
        return {"group_launches": i.group_launches, "group_jobs": i.group_jobs,
                "direct_jobs": i.direct_jobs}

    def close(self):
        if self.handle and self.ctx.handle:
            self.lib.cp_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(ctx: Context, jobs):
    arr = _job_array(list(jobs))
    _check(ctx.lib.cp_batch_run(ctx.handle, arr, len(arr), 0))


TILE_BUCKETS = (16, 32, 64, 128, 256, 512, 768, 1024, 1280)


def _encode_shape(sh):
    if sh is None:
        return 0
    g, f, x = sh[:3]
    kind = sh[3] if len(sh) > 3 else 0
    return (kind << 24) | (x << 16) | (g << 8) | f


def tile_shapes_set(ctx: Context, bytes_per_sample: int, channels: int, shapes, rel_shapes=None):
    """cp_tile_shapes_set: pins the kernel choice per frame bucket (TILE_BUCKETS) for absolute
    thresholds (`shapes`) and relative ones (`rel_shapes`, default: the same); each entry is
    (lanes_per_pixel, frames_per_lane, pixels_per_lane[, kernel]) as autotune returns it
    (kernel 1: the 8-bit packed-byte kernel with the byte search, 2: the same with a sorting
    network for the ranks, 0 or absent: the 16-bit-halves kernel), or None for the bucket's
    default.  Missing trailing buckets take their defaults."""
    nb = len(TILE_BUCKETS)
    shapes = list(shapes) + [None] * (nb - len(shapes))
    rel = shapes if rel_shapes is None else list(rel_shapes) + [None] * (nb - len(rel_shapes))
    arr = (C.c_int32 * (2 * nb))(*[_encode_shape(sh) for sh in shapes + rel])
    _check(ctx.lib.cp_tile_shapes_set(ctx.handle, bytes_per_sample, channels, arr))


def autotune(ctx: Context, bytes_per_sample: int, channels: int):
    """Runs the kernel autotuner; returns {"abs": [...], "rel": [...]}: per bucket of
    TILE_BUCKETS and threshold kind, the chosen (lanes_per_pixel, frames_per_lane,
    pixels_per_lane, kernel), or None where no register-resident kernel covers the bucket."""
    nb = len(TILE_BUCKETS)
    out = (C.c_int32 * (2 * nb))()
    _check(ctx.lib.cp_autotune(ctx.handle, bytes_per_sample, channels, out))
    dec = [None if v == 0 else ((v >> 8) & 0xFF, v & 0xFF, (v >> 16) & 0xFF, v >> 24) for v in out]
    return {"abs": dec[:nb], "rel": dec[nb:]}


# -------------------------------------------------------------------- serving

class Server:
    """cp_server: continuous-batching request loop (HBM-budget admission, stream slots)."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_inflight: int = 4):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = None
        self._keep = {}
        h = C.c_void_p()
        _check(self.lib.cp_server_create(ctx.handle, hbm_budget, max_inflight, C.byref(h)))
        self.handle = h
        ctx._children.add(self)

    def submit(self, job: Job, host_frames=None, host_out=None, host_mask=None) -> int:
        """Queues `job`; with host_frames ([n][h][w][c] array) the server uploads them into
        its own device memory.  host_out / host_mask receive the image / blend mask."""
        r = cp_request()
        r.job = job.c
        if host_frames is not None:
            host_frames = np.ascontiguousarray(host_frames)
            n, hh, ww, cc = host_frames.shape
            r.host_frames = host_frames.ctypes.data
            r.width, r.height, r.channels = ww, hh, cc
            r.bytes_per_sample = host_frames.dtype.itemsize
            r.n_frames = n
        for a in (host_out, host_mask):
            if a is not None and not a.flags.c_contiguous:
                raise ValueError("host_out / host_mask must be C-contiguous [h][w][c] arrays")
        r.host_out_px = None if host_out is None else host_out.ctypes.data
        r.host_out_mask = None if host_mask is None else host_mask.ctypes.data
        t = C.c_uint64()
        _check(self.lib.cp_server_submit(self.handle, C.byref(r), C.byref(t)))
        self._keep[t.value] = (job, host_frames, host_out, host_mask, r)
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int32()
        _check(self.lib.cp_server_poll(self.handle, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int) -> float:
        """Blocks until the request is done; returns its latency in ms."""
        ms = C.c_double()
        rc = self.lib.cp_server_wait(self.handle, ticket, C.byref(ms))
        self._keep.pop(ticket, None)
        _check(rc)
        return ms.value

    def stats(self) -> dict:
        st = cp_server_stats()
        _check(self.lib.cp_server_stats_get(self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in cp_server_stats._fields_}

    def close(self):
        if self.handle and self.ctx.handle:
            self.lib.cp_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
