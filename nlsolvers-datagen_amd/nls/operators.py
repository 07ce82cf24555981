"""Operator assembly through libnls (common/include/laplacians.hpp)."""
import ctypes as C

import numpy as np

from ._lib import as_f64, call, dptr, i32ptr, load


def laplacian_nnz(dim, n, q6=True):
    return int(load().nls_laplacian_nnz(dim, n, int(q6)))


def build_anisotropic_laplacian_noflux(nx, ny, dx, dy, c, negate=False, shift=0.0):
    """CSR (row_ptr, col, val) of div(c grad) on the (nx+2) x (ny+2) full array.

    Same arguments as the reference builder (laplacians.hpp:54-103): nx, ny are
    interior sizes, c has (nx+2)^2 entries.  negate / shift give -L + shift I.
    """
    assert nx == ny and abs(dx - dy) < 1e-10
    n = nx + 2
    c = as_f64(c).ravel()
    assert c.size == n * n
    nnz = laplacian_nnz(2, n)
    rp = np.empty(n * n + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    got = C.c_int64()
    call("nls_build_laplacian_2d", n, dx, dy, dptr(c), int(negate), float(shift), i32ptr(rp),
         i32ptr(ci), dptr(v), C.byref(got))
    return rp, ci, v


def build_anisotropic_laplacian_noflux_3d(nx, ny, nz, dx, dy, dz, c, negate=False, shift=0.0,
                                          q6=True):
    """7-point version (laplacians.hpp:158-218); q6 keeps the cross-plane couplings."""
    assert nx == ny == nz
    n = nx + 2
    c = as_f64(c).ravel()
    assert c.size == n ** 3
    nnz = laplacian_nnz(3, n, q6)
    rp = np.empty(n ** 3 + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    got = C.c_int64()
    call("nls_build_laplacian_3d", n, dx, dptr(c), int(negate), float(shift), int(q6), i32ptr(rp),
         i32ptr(ci), dptr(v), C.byref(got))
    return rp, ci, v
