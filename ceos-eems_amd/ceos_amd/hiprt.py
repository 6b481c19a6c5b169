"""Minimal ctypes access to the HIP runtime that libceos_gpu.so itself links (system ROCm).

Device buffers, streams and events for callers of the device-resident entry points
(ceos_evaluate_device).  Using the same runtime as the library matters: a second HIP runtime
in the process (for example the one bundled with a Python framework) does not share the
device with it.
"""
import ctypes as C

import numpy as np

from . import gpu

H2D, D2H, D2D = 1, 2, 3
_hip = None


def hip():
    global _hip
    if _hip is None:
        gpu.load_library()  # resolves libamdhip64.so.7 the way the library does
        L = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        vp = C.c_void_p
        L.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        L.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        L.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreateWithFlags.argtypes = [C.POINTER(vp), C.c_uint]
        L.hipStreamDestroy.argtypes = [vp]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [C.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipGetErrorString.argtypes = [C.c_int]
        L.hipStreamBeginCapture.argtypes = [vp, C.c_int]
        L.hipStreamEndCapture.argtypes = [vp, C.POINTER(vp)]
        L.hipGraphInstantiate.argtypes = [C.POINTER(vp), vp, vp, vp, C.c_size_t]
        L.hipGraphLaunch.argtypes = [vp, vp]
        L.hipGraphExecDestroy.argtypes = [vp]
        L.hipGraphDestroy.argtypes = [vp]
        _hip = L
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(rc).decode()))


def device_count():
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceArray:
    """A device buffer with a numpy dtype and shape (1-D)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def view(self, offset, n):
        """Non-owning view of elements [offset, offset + n)."""
        assert 0 <= offset and offset + n <= self.n
        return DeviceView(self.ptr + offset * self.dtype.itemsize, n, self.dtype, self)

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A window into a DeviceArray (keeps the owner alive; never frees)."""

    def __init__(self, ptr, n, dtype, owner):
        self.ptr, self.n, self.dtype, self._owner = ptr, n, np.dtype(dtype), owner
        self.nbytes = self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(hip().hipStreamCreateWithFlags(C.byref(s), 1), "hipStreamCreateWithFlags")  # non-blocking
        self.ptr = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.ptr:
                hip().hipStreamDestroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream):
        check(hip().hipEventRecord(self.ptr, stream.ptr if isinstance(stream, Stream) else stream),
              "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(ms), self.ptr, end.ptr), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                hip().hipEventDestroy(self.ptr)
        except Exception:
            pass


CAPTURE_THREAD_LOCAL = 1


class Graph:
    """A HIP graph captured from the work `fn` enqueues on `stream` (fn must only enqueue: kernels,
    device-to-device copies, event records; no memsets, no host synchronisation), replayed by
    launch().  Replays see the same device pointers, so inputs are updated in place."""

    def __init__(self, stream, fn):
        L = hip()
        check(L.hipStreamBeginCapture(stream.ptr, CAPTURE_THREAD_LOCAL), "hipStreamBeginCapture")
        g = C.c_void_p()
        try:
            fn()
        finally:
            rc = L.hipStreamEndCapture(stream.ptr, C.byref(g))
        check(rc, "hipStreamEndCapture")
        e = C.c_void_p()
        check(L.hipGraphInstantiate(C.byref(e), g, None, None, 0), "hipGraphInstantiate")
        self.graph, self.exec = g.value, e.value

    def launch(self, stream):
        check(hip().hipGraphLaunch(self.exec, stream.ptr), "hipGraphLaunch")

    def __del__(self):
        try:
            hip().hipGraphExecDestroy(self.exec)
            hip().hipGraphDestroy(self.graph)
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
