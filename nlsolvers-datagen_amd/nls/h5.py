"""HDF5 trajectory files in the launchers' per-run layout, written by libnls.

Replaces Launcher.save_to_hdf5 (scripts/real_launcher_2d.py:201-245,
scripts/real_launcher_3d.py:206-252, scripts/complex_launcher_2d.py:193-236) and the
h5py dependency: libnls loads libhdf5 at run time and writes the same groups,
attributes and datasets with the types h5py uses (variable-length UTF-8 strings,
int64, float64, complex128 as a compound {r, i}).  `AsyncWriter` writes files from a
background thread so the next batch of trajectories runs while the last is saved.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import call, dptr


def available():
    """libhdf5 version string, or None when libhdf5 cannot be loaded."""
    buf = C.create_string_buffer(32)
    rc = L.load().nls_h5_available(buf, 32)
    return buf.value.decode() if rc == 0 else None


def _b(s):
    return None if s is None else str(s).encode()


class _Run:
    """Keeps the ctypes record and every buffer it points to alive."""

    def __init__(self, path, *, u0, m, c, u, v=None, v0=None, Lx, Ly=None, Lz=None, T, nt,
                 num_snapshots, problem_type, run_id, run_index, phenomenon="",
                 phenomenon_params=None, m_type="constant", elapsed_time=0.0, timestamp=None,
                 c_type=""):
        u0 = np.asarray(u0)
        cplx = np.iscomplexobj(u0)
        dim = u0.ndim
        if dim not in (2, 3):
            raise ValueError("u0 must be 2D or 3D")
        grid = u0.shape
        keep = []

        def arr(a, complex_ok=False):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.complex128 if (complex_ok and cplx) else np.float64)
            keep.append(a)
            return a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))

        u = np.asarray(u)
        if u.ndim != dim + 1:
            raise ValueError("u must be (num_snapshots_stored, *grid)")
        if not cplx and (v is None or v0 is None):
            raise ValueError("real runs need v0 and v")
        for name, a in (("m", m), ("c", c)) + ((("v0", v0),) if not cplx else ()):
            if np.shape(a) != grid:
                raise ValueError(f"{name} must have the grid shape {grid}")
        params = phenomenon_params or {}
        self._keys = [_b(k) for k in params]
        self._vals = [_b(str(v_)) for v_ in params.values()]
        kt = (C.c_char_p * max(len(params), 1))(*self._keys)
        vt = (C.c_char_p * max(len(params), 1))(*self._vals)
        keep += [kt, vt]
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        self.rec = L.H5Run(
            _b(path), int(cplx), dim, grid[0], grid[1], grid[2] if dim == 3 else 1,
            float(Lx), float(Ly), float(Lz), float(T), int(nt), int(num_snapshots),
            _b(problem_type), _b(run_id), _b(phenomenon), _b(m_type), _b(timestamp),
            int(run_index), float(elapsed_time), len(params), kt, vt,
            arr(u0, True), None if cplx else arr(v0), arr(m), arr(c),
            u.shape[0], u.shape[1], u.shape[2], u.shape[3] if dim == 3 else 1,
            arr(u, True), None if cplx else arr(v), _b(c_type) if c_type else None)
        self._keep = keep


def save_run(path, **kw):
    """Write one trajectory file now.  Keyword arguments: u0, m, c, u (stored
    trajectory, complex for NLSE runs), v0 and v for real runs, Lx[, Ly, Lz], T, nt,
    num_snapshots, problem_type, run_id, run_index, phenomenon, phenomenon_params
    (dict, stored as metadata attributes phenomenon_<key> = str(value)), m_type, c_type
    (metadata attribute c_type when set), elapsed_time, timestamp (default: now)."""
    r = _Run(path, **kw)
    call("nls_h5_write_run", C.byref(r.rec))
    return path


class AsyncWriter:
    """Background HDF5 writer: submit() copies the arrays and returns; files are
    written in submission order; drain() waits and raises the first failure."""

    def __init__(self, max_queued_bytes=1 << 30):
        self.h = C.c_void_p()
        call("nls_h5_writer_create", C.c_size_t(max_queued_bytes), C.byref(self.h))

    def submit(self, path, **kw):
        r = _Run(path, **kw)
        call("nls_h5_writer_submit", self.h, C.byref(r.rec))
        return path

    def drain(self):
        n = C.c_int64()
        call("nls_h5_writer_drain", self.h, C.byref(n))
        return n.value

    def close(self):
        if self.h and self.h.value:
            try:
                self.drain()
            finally:
                call("nls_h5_writer_destroy", self.h)
                self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read(path, name):
    """Dataset `name` as a numpy array (complex128 for compound {r, i} data)."""
    dims = (C.c_int64 * 4)()
    nd = C.c_int()
    cx = C.c_int()
    call("nls_h5_read_dataset", _b(path), _b(name), None, 0, dims, C.byref(nd), C.byref(cx))
    shape = tuple(dims[k] for k in range(nd.value))
    count = int(np.prod(shape)) if shape else 1
    buf = np.empty(count * (2 if cx.value else 1))
    call("nls_h5_read_dataset", _b(path), _b(name), dptr(buf), buf.size, dims, C.byref(nd),
         C.byref(cx))
    return (buf.view(np.complex128) if cx.value else buf).reshape(shape)


def shape(path, name):
    """(shape, is_complex) of dataset `name` without reading it."""
    dims = (C.c_int64 * 4)()
    nd = C.c_int()
    cx = C.c_int()
    call("nls_h5_read_dataset", _b(path), _b(name), None, 0, dims, C.byref(nd), C.byref(cx))
    return tuple(dims[k] for k in range(nd.value)), bool(cx.value)


def read_attr(path, obj, name):
    """Attribute `name` of group `obj`: str or float."""
    num = C.c_double()
    s = C.create_string_buffer(4096)
    is_s = C.c_int()
    call("nls_h5_read_attr", _b(path), _b(obj), _b(name), C.byref(num), s, 4096, C.byref(is_s))
    return s.value.decode() if is_s.value else num.value
