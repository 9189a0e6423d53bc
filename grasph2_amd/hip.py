"""Minimal ctypes binding of the HIP runtime (device memory, streams, events).

Plumbing for the tests and the benchmark: it lets Python hand device pointers
to the C ABI of libgrasph2_amd.so without importing a framework.
"""
import ctypes
import os

_HIP = None

H2D, D2H, D2D = 1, 2, 3


def _lib():
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64.so not found")
        h = _HIP
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return _HIP


def check(code, what="HIP call"):
    if code != 0:
        raise RuntimeError("%s failed: %s (%d)" % (what, _lib().hipGetErrorString(code).decode(), code))


def device_count():
    n = ctypes.c_int(0)
    code = _lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if code == 0 else 0


def set_device(d):
    check(_lib().hipSetDevice(d), "hipSetDevice")


def synchronize():
    check(_lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    """(free, total) bytes of device memory on the current device."""
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(_lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """Owning device allocation of `nbytes` bytes."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(_lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc(%d)" % self.nbytes)

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        check(_lib().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(_lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        check(_lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def copy_from(self, other):
        check(_lib().hipMemcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), D2D), "hipMemcpy D2D")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _lib().hipFree(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(_lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(_lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle.value:
                _lib().hipStreamDestroy(self.handle)
        except Exception:
            pass


class Graph:
    """Work enqueued on a stream by fn(stream), captured (hipStreamBeginCapture,
    global mode) into an executable HIP graph; launch() replays it."""

    def __init__(self, stream, fn):
        h = _lib()
        h.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        h.hipGraphInstantiate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_size_t]
        h.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
        h.hipGraphDestroy.argtypes = [ctypes.c_void_p]
        self.graph, self.exec = ctypes.c_void_p(), ctypes.c_void_p()
        check(h.hipStreamBeginCapture(stream.handle, 0), "hipStreamBeginCapture")
        try:
            fn(stream)
        finally:
            code = h.hipStreamEndCapture(stream.handle, ctypes.byref(self.graph))
        check(code, "hipStreamEndCapture")
        check(h.hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, 0), "hipGraphInstantiate")

    def launch(self, stream):
        check(_lib().hipGraphLaunch(self.exec, stream.handle), "hipGraphLaunch")

    def __del__(self):
        try:
            if self.exec.value:
                _lib().hipGraphExecDestroy(self.exec)
            if self.graph.value:
                _lib().hipGraphDestroy(self.graph)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(_lib().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        check(_lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(_lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = ctypes.c_float(0)
        check(_lib().hipEventElapsedTime(ctypes.byref(t), self.handle, end.handle), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        try:
            if self.handle.value:
                _lib().hipEventDestroy(self.handle)
        except Exception:
            pass
