"""Minimal HIP runtime plumbing through ctypes: device buffers and copies for
callers that hand device-resident filterbanks to hd_execute_device."""
import ctypes as C
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("HIP runtime (libamdhip64) not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


def device_count():
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        check(hip().hipMalloc(C.byref(self.ptr), max(nbytes, 1)), "hipMalloc")

    def upload(self, array):
        import numpy as np
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, nbytes=None):
        """Bytes of the buffer as a uint8 numpy array."""
        import numpy as np
        n = self.nbytes if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, n, D2H), "hipMemcpy D2H")
        return out

    def upload_at(self, offset, array):
        """Copies a host array to the buffer starting `offset` bytes in."""
        import numpy as np
        a = np.ascontiguousarray(array)
        assert 0 <= offset and offset + a.nbytes <= self.nbytes
        check(hip().hipMemcpy(C.c_void_p(self.ptr.value + offset), a.ctypes.data, a.nbytes, H2D),
              "hipMemcpy H2D")

    def download_at(self, offset, nbytes):
        """nbytes of the buffer from `offset` on, as a uint8 numpy array."""
        import numpy as np
        assert 0 <= offset and offset + nbytes <= self.nbytes
        out = np.empty(nbytes, np.uint8)
        check(hip().hipMemcpy(out.ctypes.data, C.c_void_p(self.ptr.value + offset), nbytes, D2H),
              "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def address(self):
        return self.ptr.value


class PinnedBuffer:
    """Page-locked host memory (hipHostMalloc): host filterbanks that the
    runtime copies to the device at full link speed."""

    def __init__(self, nbytes):
        h = hip()
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        check(h.hipHostMalloc(C.byref(self.ptr), max(nbytes, 1), 0), "hipHostMalloc")

    def array(self):
        """The buffer as a uint8 numpy array (no copy)."""
        import numpy as np
        return np.ctypeslib.as_array(C.cast(self.ptr, C.POINTER(C.c_uint8)), (self.nbytes,))

    def free(self):
        if self.ptr:
            hip().hipHostFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def address(self):
        return self.ptr.value


def copy_device(dst, src, nbytes):
    """Device-to-device copy (one runtime copy kernel of nbytes)."""
    check(hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, D2D), "hipMemcpy D2D")


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream at the default priority, for copies that run beside the
    pipeline's own stream."""

    def __init__(self):
        h = hip()
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        self.ptr = C.c_void_p()
        check(h.hipStreamCreate(C.byref(self.ptr)), "hipStreamCreate")

    def copy(self, dst, src, nbytes, kind):
        """Asynchronous copy of nbytes (addresses as integers) on this stream."""
        check(hip().hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(src), nbytes, kind, self.ptr),
              "hipMemcpyAsync")

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def destroy(self):
        if self.ptr:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
