"""Minimal ctypes binding of the HIP runtime used as plumbing by the Python
front end, the tests and the benchmark: device buffers, copies, streams,
events and graphs.  No torch is involved, so exactly one HIP runtime
(``libamdhip64.so.7`` from /opt/rocm) is loaded into the process."""
import ctypes
import os

import numpy as np

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3
hipStreamCaptureModeGlobal = 0
hipStreamCaptureModeThreadLocal = 1

ERROR_NAMES = {
    0: "hipSuccess", 1: "hipErrorInvalidValue", 2: "hipErrorOutOfMemory",
    9: "hipErrorInvalidConfiguration", 17: "hipErrorInvalidDevicePointer",
    100: "hipErrorNoDevice", 101: "hipErrorInvalidDevice",
}
hipSuccess = 0
hipErrorInvalidValue = 1
hipErrorInvalidConfiguration = 9
hipErrorInvalidDevicePointer = 17


class HipError(RuntimeError):
    def __init__(self, code, what=""):
        name = ERROR_NAMES.get(code, str(code))
        try:
            msg = runtime().hipGetErrorString(code).decode()
        except Exception:  # noqa: BLE001 - best effort message
            msg = ""
        super().__init__(f"{what}: {name} ({code}) {msg}".strip())
        self.code = code


def _candidates():
    # by soname first: if another component of the process (e.g. torch) has
    # already loaded a HIP runtime, this binds to that same copy
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return ["libamdhip64.so.7", os.path.join(rocm, "lib", "libamdhip64.so.7"),
            os.path.join(rocm, "lib", "libamdhip64.so")]


def runtime():
    """The loaded HIP runtime (RTLD_GLOBAL so libpbe_hip.so binds to it)."""
    global _hip
    if _hip is None:
        err = None
        for path in _candidates():
            try:
                _hip = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
                break
            except OSError as e:  # try the next candidate
                err = e
        if _hip is None:
            raise OSError(f"cannot load the HIP runtime: {err}")
        h = _hip
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [i]
        for name, args in [
            ("hipGetDeviceCount", [ctypes.POINTER(i)]),
            ("hipSetDevice", [i]),
            ("hipMalloc", [ctypes.POINTER(vp), sz]),
            ("hipFree", [vp]),
            ("hipMemcpy", [vp, vp, sz, i]),
            ("hipMemcpyAsync", [vp, vp, sz, i, vp]),
            ("hipMemset", [vp, i, sz]),
            ("hipMemsetAsync", [vp, i, sz, vp]),
            ("hipDeviceSynchronize", []),
            ("hipStreamCreate", [ctypes.POINTER(vp)]),
            ("hipStreamDestroy", [vp]),
            ("hipStreamSynchronize", [vp]),
            ("hipEventCreate", [ctypes.POINTER(vp)]),
            ("hipEventDestroy", [vp]),
            ("hipEventRecord", [vp, vp]),
            ("hipEventSynchronize", [vp]),
            ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), vp, vp]),
            ("hipStreamBeginCapture", [vp, i]),
            ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
            ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, vp, sz]),
            ("hipGraphLaunch", [vp, vp]),
            ("hipGraphExecDestroy", [vp]),
            ("hipGraphDestroy", [vp]),
            ("hipGetLastError", []),
            ("hipPeekAtLastError", []),
        ]:
            f = getattr(h, name)
            f.restype = i
            f.argtypes = args
    return _hip


def check(code, what="hip"):
    if code != 0:
        raise HipError(code, what)


def device_count():
    n = ctypes.c_int(0)
    code = runtime().hipGetDeviceCount(ctypes.byref(n))
    return n.value if code == 0 else 0


def synchronize():
    check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A float64 (or raw byte) device buffer owned by Python."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 8)
        p = ctypes.c_void_p()
        check(runtime().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        check(runtime().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                  a.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")

    def download(self, count=None):
        count = self.n if count is None else count
        out = np.empty(count, dtype=self.dtype)
        check(runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                  out.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    def zero(self, stream=None):
        if stream is None:
            check(runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        else:
            check(runtime().hipMemsetAsync(self.ptr, 0, self.nbytes, stream.handle),
                  "hipMemsetAsync")

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr is not None and self.ptr.value:
            runtime().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(runtime().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s

    def synchronize(self):
        check(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                runtime().hipStreamDestroy(self.handle)
        except Exception:  # noqa: BLE001
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e

    def record(self, stream):
        check(runtime().hipEventRecord(self.handle, stream.handle if stream else None),
              "hipEventRecord")

    def synchronize(self):
        check(runtime().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float(0.0)
        check(runtime().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle),
              "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                runtime().hipEventDestroy(self.handle)
        except Exception:  # noqa: BLE001
            pass
