"""Minimal HIP runtime access (ctypes) for benchmarks and tests: device buffers, events and
copies on the library's own stream.  Uses the same libamdhip64 that libchrono_hip.so is
linked against, so only one HIP runtime lives in the process."""
import ctypes as C

from . import load_library

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        load_library()  # brings in the runtime libchrono_hip.so was linked with
        for name in ("libamdhip64.so.7", "libamdhip64.so.6", "libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64 not found")
        h = _hip
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy2D.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                  C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostFree.argtypes = [C.c_void_p]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
    return _hip


def check(rc: int, what: str = "hip call"):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def to_host(self, out):
        """Copies the first out.nbytes bytes into the numpy array `out`."""
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, hipMemcpyDeviceToHost), "hipMemcpy")
        return out

    def from_host(self, arr):
        check(hip().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, hipMemcpyHostToDevice), "hipMemcpy")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream):
        check(hip().hipEventRecord(self.e, C.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value


def copy_rows_to_host(out, dev_ptr: int, frame_bytes: int, n_frames: int):
    """out[f] = the first out[f].nbytes bytes of device frame f (a band of leading rows)."""
    width = out.nbytes // n_frames
    check(hip().hipMemcpy2D(out.ctypes.data, width, dev_ptr, frame_bytes, width, n_frames,
                            hipMemcpyDeviceToHost), "hipMemcpy2D")
    return out


class PinnedArray:
    """numpy view of page-locked host memory (hipHostMalloc), for asynchronous copies."""

    def __init__(self, shape, dtype, flags: int = 0):
        import numpy as np
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        check(hip().hipHostMalloc(C.byref(p), max(nbytes, 1), flags), "hipHostMalloc")
        self.ptr = p.value
        buf = (C.c_uint8 * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            hip().hipHostFree(self.ptr)
            self.ptr = None
