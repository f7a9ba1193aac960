"""Device context and arrays over the C ABI (no torch: the library owns its HIP stream,
memory, events and graphs)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B

_default = None


class Context:
    """One HIP stream on one device (bb_ctx)."""

    def __init__(self, device: int = 0):
        L = B.lib()
        h = C.c_void_p()
        rc = L.bb_ctx_create(device, C.byref(h))
        if rc != B.BB_OK:
            raise B.BBError(rc, f"cannot create a context on device {device} (is a GPU visible?)")
        self.h = h
        self.device = device
        B.track(self, context=True)

    def close(self):
        if self.h:
            B.lib().bb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        B.check(rc, self.h)

    def sync(self):
        self.check(B.lib().bb_synchronize(self.h))

    def info(self):
        arch = C.create_string_buffer(64)
        name = C.create_string_buffer(128)
        cu = C.c_int()
        self.check(B.lib().bb_device_info(self.h, arch, 64, name, 128, C.byref(cu)))
        return {"arch": arch.value.decode(), "name": name.value.decode(), "cu": cu.value}

    def timer_start(self):
        self.check(B.lib().bb_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        self.check(B.lib().bb_timer_stop(self.h, C.byref(ms)))
        return ms.value


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


_DT = {np.dtype(np.float64): B.BB_F64, np.dtype(np.float32): B.BB_F32}
# dtypes a device array may hold (floating point data and the integer index tables)
STORABLE = {np.dtype(t) for t in (np.float64, np.float32, np.int64, np.int32, np.int16, np.int8, np.uint8)}


def host_buffer(a):
    """(contiguous 1-D buffer in memory order, dtype, strides in elements) for an upload.
    The dtype is kept as given; only non-storable dtypes are refused."""
    a = np.asarray(a)
    if a.dtype not in STORABLE:
        raise TypeError(f"cannot store dtype {a.dtype} on the device")
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        a = np.ascontiguousarray(a)
    strides = tuple(s // a.itemsize for s in a.strides)
    return np.ascontiguousarray(a.ravel(order="K")), a.dtype, strides, a.shape


class DeviceArray:
    """A strided device array.  ``strides`` are in elements (like Julia, unlike numpy)."""

    def __init__(self, shape, dtype, strides=None, ctx: Context | None = None, _ptr=None, _owner=None):
        self.ctx = ctx or default_context()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if strides is None:
            strides, acc = [], 1
            for s in reversed(self.shape):
                strides.insert(0, acc)
                acc *= s
        self.strides = tuple(int(s) for s in strides)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes_span = self._span() * self.dtype.itemsize
        if _ptr is None:
            p = C.c_void_p()
            self.ctx.check(B.lib().bb_malloc(self.ctx.h, max(self.nbytes_span, 1), C.byref(p)))
            self.ptr = p.value
            self._owner = None
            self._owns = True
            B.track(self)
        else:
            self.ptr = _ptr
            self._owner = _owner
            self._owns = False

    def _span(self):
        if any(s == 0 for s in self.shape):
            return 0
        return 1 + sum((n - 1) * st for n, st in zip(self.shape, self.strides))

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def code(self):
        if self.dtype not in _DT:
            raise TypeError(f"operators take float64 or float32 data, not {self.dtype}")
        return _DT[self.dtype]

    def free(self):
        if self._owns and self.ptr:
            B.lib().bb_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def from_numpy(cls, a, ctx: Context | None = None):
        buf, dtype, strides, shape = host_buffer(a)
        d = cls(shape, dtype, strides, ctx)
        d.ctx.check(B.lib().bb_memcpy_h2d(d.ctx.h, C.c_void_p(d.ptr), buf.ctypes.data_as(C.c_void_p), buf.nbytes))
        return d

    @classmethod
    def empty(cls, shape, dtype=np.float64, order="C", ctx: Context | None = None):
        shape = tuple(shape)
        if order == "F":
            strides, acc = [], 1
            for s in shape:
                strides.append(acc)
                acc *= s
            return cls(shape, dtype, strides, ctx)
        return cls(shape, dtype, None, ctx)

    @classmethod
    def zeros(cls, shape, dtype=np.float64, order="C", ctx: Context | None = None):
        d = cls.empty(shape, dtype, order, ctx)
        d.ctx.check(B.lib().bb_fill_zero(d.ctx.h, C.c_void_p(d.ptr), d.nbytes_span))
        return d

    def numpy(self) -> np.ndarray:
        span = self._span()
        buf = np.empty(span, self.dtype)
        self.ctx.check(B.lib().bb_memcpy_d2h(self.ctx.h, buf.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                             buf.nbytes))
        return np.lib.stride_tricks.as_strided(buf, self.shape, [s * self.dtype.itemsize for s in self.strides]).copy()

    def count_nonfinite(self) -> int:
        """number of NaN / Inf values in the array's memory span (a device scan)"""
        n = C.c_int64()
        rc = B.lib().bb_count_nonfinite(self.ctx.h, self.code, C.c_void_p(self.ptr), self._span(), C.byref(n))
        if rc not in (B.BB_OK, B.BB_ERR_NONFINITE):
            self.ctx.check(rc)
        return n.value

    def view(self, shape, strides) -> "DeviceArray":
        """A re-strided view of the same memory (no copy, no bounds growth)."""
        v = DeviceArray(shape, self.dtype, strides, self.ctx, _ptr=self.ptr, _owner=self)
        if v._span() > self._span():
            raise ValueError("view exceeds the allocation")
        return v

    @property
    def T(self):
        return self.view(self.shape[::-1], self.strides[::-1])
