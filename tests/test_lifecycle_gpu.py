"""Handle lifetimes through the C ABI: a context destroyed before its stacks, batches and
servers releases their device memory and leaves inert handles whose own destroy calls are
safe; the server refuses malformed fade tables before anything reaches the device."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import chrono_photo_amd as cpa

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Runs in its own process: a use-after-free here would end that process, not the test run.
CTX_FIRST = r'''
import ctypes as C, sys
import numpy as np
sys.path[:0] = [sys.argv[1]]
import chrono_photo_amd as cpa
print("case: context destroyed before a stack, a borrowed stack, a batch and a server", flush=True)
lib = cpa.load_library()
ctx = C.c_void_p()
assert lib.cp_ctx_create(0, 0, C.byref(ctx)) == 0
st = C.c_void_p()
assert lib.cp_stack_create(ctx, 64, 32, 3, 1, 8, C.byref(st)) == 0
frames = np.random.default_rng(1).integers(0, 256, (8, 32, 64, 3), np.uint8)
assert lib.cp_stack_upload(st, 0, frames.ctypes.data, 8) == 0
info = cpa.cp_stack_info()
assert lib.cp_stack_info_get(st, C.byref(info)) == 0
wrapped = C.c_void_p()
assert lib.cp_stack_wrap(ctx, info.device_ptr, 64, 32, 3, 1, 4, C.byref(wrapped)) == 0
from chrono_photo_amd import hip as H
px = H.DeviceBuffer(64 * 32 * 3)
proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), cpa.BackgroundMode.FIRST,
                            cpa.OutlierSelectionMode.EXTREME)
job = cpa.cp_job()
job.mode = cpa.CP_MODE_OUTLIER
job.stack = st
job.frames = cpa.cp_frames(None, 8, 0, None, 0, 0)
job.outlier = proc.params(1)
job.out = cpa.cp_outputs(px.ptr, None, None, None)
jobs = (cpa.cp_job * 1)(job)
batch = C.c_void_p()
assert lib.cp_batch_create(ctx, jobs, 1, 0, C.byref(batch)) == 0
assert lib.cp_batch_launch(batch) == 0
srv = C.c_void_p()
assert lib.cp_server_create(ctx, 0, 2, C.byref(srv)) == 0
req = cpa.cp_request()
req.job = job
host_out = np.zeros((32, 64, 3), np.uint8)
req.host_out_px = host_out.ctypes.data
t = C.c_uint64()
assert lib.cp_server_submit(srv, C.byref(req), C.byref(t)) == 0

assert lib.cp_ctx_destroy(ctx) == 0                      # the context goes first
assert lib.cp_stack_upload(st, 0, frames.ctypes.data, 1) == cpa.CP_ERR_INVALID_ARG
assert lib.cp_batch_launch(batch) == cpa.CP_ERR_INVALID_ARG
assert lib.cp_server_submit(srv, C.byref(req), C.byref(t)) == cpa.CP_ERR_INVALID_ARG
assert lib.cp_server_wait(srv, t.value, None) == 0        # drained before the context went
assert lib.cp_stack_destroy(wrapped) == 0
assert lib.cp_stack_destroy(st) == 0
assert lib.cp_batch_destroy(batch) == 0
assert lib.cp_server_destroy(srv) == 0
print("LIFECYCLE_OK", flush=True)
'''


def test_context_destroyed_before_its_handles():
    r = subprocess.run([sys.executable, "-c", CTX_FIRST, os.path.join(ROOT, "chrono-photo_amd")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "LIFECYCLE_OK" in r.stdout


@pytest.mark.parametrize("mode", [cpa.CP_MODE_DARKER, cpa.CP_MODE_LIGHTER, cpa.CP_MODE_OUTLIER])
@pytest.mark.parametrize("bad", ["empty", "null"])
def test_server_refuses_malformed_fade(ctx, mode, bad):
    x = cpa.Scene(seed=3, width=32, height=16, n_frames=6).generate()
    st = cpa.Stack.from_array(ctx, x)
    proc = None
    if mode == cpa.CP_MODE_OUTLIER:
        proc = cpa.OutlierProcessor(cpa.Threshold.abs(0.05, 0.2), cpa.BackgroundMode.FIRST,
                                    cpa.OutlierSelectionMode.FORWARD)
    job = cpa.Job(mode, st, None, processor=proc)
    table = np.ones(4, np.float32)
    f = cpa.cp_fade(0, 0, 1, 0, 0 if bad == "empty" else 4,
                    None if bad == "null" else table.ctypes.data_as(C.POINTER(C.c_float)))
    if mode == cpa.CP_MODE_OUTLIER:
        job.c.outlier.fade = f
    else:
        job.c.fade = f
    srv = cpa.Server(ctx, max_inflight=2)
    req = cpa.cp_request()
    req.job = job.c
    out = np.zeros((16, 32, 3), np.uint8)
    req.host_out_px = out.ctypes.data
    t = C.c_uint64()
    assert ctx.lib.cp_server_submit(srv.handle, C.byref(req), C.byref(t)) == cpa.CP_ERR_INVALID_ARG
    assert b"fade" in ctx.lib.cp_last_error()
    assert srv.stats()["submitted"] == 0
    srv.close()
    st.close()
