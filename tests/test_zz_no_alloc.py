"""cp_shake_noalloc (the shake search with its scratch in a static device array) against
cp_shake through the C ABI: identical offsets on consecutive calls that share the scratch,
the second smaller than the first, with host and device outputs; oversized requests are
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import chrono_photo_amd as cpa
from chrono_photo_amd import hip as H

pytestmark = pytest.mark.gpu


def _stack(ctx, seed, n, w, h, bps, shake_max):
    sc = cpa.Scene(seed=seed, width=w, height=h, bytes_per_sample=bps, n_frames=n,
                   base_lo=8000 if bps == 2 else 40, base_hi=50000 if bps == 2 else 200,
                   sigma=200 if bps == 2 else 3, n_disks=1, shake_max=shake_max,
                   texture_amp=3000 if bps == 2 else 30)
    return cpa.Stack.from_array(ctx, sc.generate())


def _plain(ctx, st, n, anchors, r, s, frames=None):
    a = np.ascontiguousarray(np.asarray(anchors, np.int32).reshape(-1))
    fr = None if frames is None else np.ascontiguousarray(frames, np.uint32)
    out = np.zeros((n, 2), np.int32)
    cpa._check(ctx.lib.cp_shake(ctx.handle, st.handle, None if fr is None else fr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                n, a.ctypes.data_as(C.POINTER(C.c_int32)), len(a) // 2, r, s,
                                out.ctypes.data_as(C.POINTER(C.c_int32))))
    return out


def _noalloc(ctx, st, n, anchors, r, s, frames=None, device_out=False):
    a = np.ascontiguousarray(np.asarray(anchors, np.int32).reshape(-1))
    fr = None if frames is None else np.ascontiguousarray(frames, np.uint32)
    fp = None if fr is None else fr.ctypes.data_as(C.POINTER(C.c_uint32))
    ap = a.ctypes.data_as(C.POINTER(C.c_int32))
    out = np.zeros((n, 2), np.int32)
    if device_out:
        buf = H.DeviceBuffer(out.nbytes)
        rc = ctx.lib.cp_shake_noalloc(ctx.handle, st.handle, fp, n, ap, len(a) // 2, r, s, C.c_void_p(buf.ptr),
                                      cpa.CP_FLAG_DEVICE_OUT)
        cpa._check(rc)
        ctx.synchronize()
        buf.to_host(out)
        buf.free()
    else:
        cpa._check(ctx.lib.cp_shake_noalloc(ctx.handle, st.handle, fp, n, ap, len(a) // 2, r, s,
                                            C.c_void_p(out.ctypes.data), 0))
    return out


@pytest.mark.parametrize("bps", [1, 2])
@pytest.mark.parametrize("device_out", [False, True])
def test_noalloc_equals_plain_on_consecutive_calls(ctx, bps, device_out):
    print(f"case: cp_shake_noalloc bps={bps} device_out={device_out}", flush=True)
    # first call: 40 frames, 3 anchors, r=6, s=5; second: 12 frames, one anchor, s=3
    st = _stack(ctx, 70 + bps, 40, 160, 120, bps, 4)
    anchors = [(40, 40), (110, 60), (80, 90)]
    want = _plain(ctx, st, 40, anchors, 6, 5)
    got = _noalloc(ctx, st, 40, anchors, 6, 5, device_out=device_out)
    assert np.array_equal(got, want)
    small = _stack(ctx, 80 + bps, 12, 96, 64, bps, 2)
    frames = [0, 2, 3, 5, 7, 8, 9, 11]
    want2 = _plain(ctx, small, len(frames), [(48, 32)], 5, 3, frames=frames)
    got2 = _noalloc(ctx, small, len(frames), [(48, 32)], 5, 3, frames=frames, device_out=device_out)
    assert np.array_equal(got2, want2)
    st.close()
    small.close()


def test_noalloc_trivial_and_oversized(ctx):
    st = _stack(ctx, 91, 6, 64, 48, 1, 2)
    # no anchors: (0, 0) then (-s, -s), like cp_shake
    for dev in (False, True):
        got = _noalloc(ctx, st, 6, np.zeros((0, 2), np.int32), 4, 3, device_out=dev)
        assert np.array_equal(got, _plain(ctx, st, 6, np.zeros((0, 2), np.int32), 4, 3))
    st.close()
    # 1,800 frames x 625 candidate offsets (s = 12) x 8 bytes exceed the 8 MiB scratch
    big = cpa.Stack.from_array(ctx, np.zeros((1800, 40, 40, 3), np.uint8))
    a = np.array([20, 20], np.int32)
    out = np.zeros((1800, 2), np.int32)
    rc = ctx.lib.cp_shake_noalloc(ctx.handle, big.handle, None, 1800, a.ctypes.data_as(C.POINTER(C.c_int32)), 1, 2,
                                  12, C.c_void_p(out.ctypes.data), 0)
    assert rc == cpa.CP_ERR_CAPACITY
    assert b"scratch" in ctx.lib.cp_last_error()
    big.close()
