"""Device buffers and streams over libphx.so's runtime plumbing (include/phx_runtime.h)."""
import ctypes as C

import numpy as np

from ._abi import PHX_OK, PhxError, lib


def _chk(rc, what):
    if rc != PHX_OK:
        raise PhxError(f"{what} failed with status {rc}")


class DeviceArray:
    """A device allocation with a numpy-like shape and dtype (host copies on request)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(lib().phx_rt_malloc(C.byref(p), max(self.nbytes, 8)), "device malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _chk(lib().phx_rt_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes), "copy to device")

    def get(self):
        out = np.empty(self.shape, self.dtype)
        _chk(lib().phx_rt_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "copy to host")
        return out

    def zero(self):
        _chk(lib().phx_rt_memset(self.ptr, 0, self.nbytes), "memset")

    def fill_from(self, other):
        assert other.nbytes == self.nbytes
        _chk(lib().phx_rt_memcpy_dtod(self.ptr, other.ptr, self.nbytes), "device copy")

    def view(self, offset, shape):
        """A non-owning view of `shape` elements starting `offset` elements in."""
        return DeviceView(self, offset, shape)

    def free(self):
        if self.ptr:
            lib().phx_rt_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView(DeviceArray):
    """Part of a DeviceArray (keeps its parent alive, frees nothing)."""

    def __init__(self, parent, offset, shape):
        self.parent = parent
        self.shape = tuple(int(s) for s in shape)
        self.dtype = parent.dtype
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        off = int(offset) * self.dtype.itemsize
        assert 0 <= off and off + self.nbytes <= parent.nbytes, "view outside its array"
        self.ptr = parent.ptr + off

    def free(self):
        self.ptr = None


def synchronize():
    _chk(lib().phx_rt_device_synchronize(), "device synchronize")


def device_count():
    n = C.c_int(0)
    if lib().phx_rt_device_count(C.byref(n)) != PHX_OK:
        return 0
    return n.value


class Event:
    def __init__(self):
        e = C.c_void_p()
        _chk(lib().phx_rt_event_create(C.byref(e)), "event create")
        self.e = e.value

    def record(self, stream=None):
        _chk(lib().phx_rt_event_record(self.e, stream), "event record")

    def elapsed_ms(self, later):
        ms = C.c_float()
        _chk(lib().phx_rt_event_elapsed_ms(self.e, later.e, C.byref(ms)), "event elapsed")
        return ms.value

    def __del__(self):
        try:
            lib().phx_rt_event_destroy(self.e)
        except Exception:
            pass
