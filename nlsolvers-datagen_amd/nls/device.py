"""Device memory through libnls (no other GPU runtime needed on the Python side)."""
import ctypes as C

import numpy as np

from ._lib import call


class DeviceArray:
    """A device allocation mirroring a numpy array's dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("nls_dev_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            call("nls_memcpy_h2d", C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes)
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("nls_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                 self.nbytes)
        return out

    def free(self):
        if self.ptr:
            call("nls_dev_free", C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceSpMV:
    """y = A x for a device CSR (DeviceSpMV<T>, device/include/spmv.hpp:7-73)."""

    _FN = {np.dtype(np.float64): "nls_csr_spmv_f64", np.dtype(np.float32): "nls_csr_spmv_f32",
           np.dtype(np.complex128): "nls_csr_spmv_c128", np.dtype(np.complex64): "nls_csr_spmv_c64"}

    def __init__(self, row_ptr, col, values):
        self.n = len(row_ptr) - 1
        self.values_dtype = np.asarray(values).dtype
        self.rp = DeviceArray.from_host(np.asarray(row_ptr, np.int32))
        self.ci = DeviceArray.from_host(np.asarray(col, np.int32))
        self.val = DeviceArray.from_host(np.asarray(values))

    def multiply(self, x, y):
        """x, y: DeviceArray of the operator's value dtype."""
        fn = self._FN[self.values_dtype]
        call(fn, C.c_int64(self.n), C.c_void_p(self.rp.ptr), C.c_void_p(self.ci.ptr),
             C.c_void_p(self.val.ptr), C.c_void_p(x.ptr), C.c_void_p(y.ptr), None)
        call("nls_dev_synchronize")
