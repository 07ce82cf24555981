"""Solver objects mirroring the reference's device classes.

Each class forwards to the C ABI of libnls; names, arguments and call order
follow the reference so that driver code reads the same:

  MatrixFunctionApplicatorReal     device/include/matfunc_real.hpp:100-419
  MatrixFunctionApplicatorComplex  device/include/matfunc_complex.hpp:112-438
  KGESolverDevice                  device/include/kg_dev.hpp:19-157
  NLSESolverDevice                 device/include/nlse_dev.hpp:66-362
  SP4SolverDevice                  device/include/stochastic_phi4_dev.hpp:17-160
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import as_c128_view, as_f64, call, dptr, i32ptr


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = C.c_void_p()

    def close(self):
        if self.h and self.h.value:
            call(self._destroy, self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MatrixFunctionApplicator(_Handle):
    """FP64 Krylov matrix function on a host CSR (reference Lanczos semantics)."""
    _destroy = "nls_matfunc_destroy"
    COS_SQRT, SINC2_SQRT, ID_SQRT, SINC2_HALF, COSM1_SQRT = (L.F_COS_SQRT, L.F_SINC2_SQRT,
                                                             L.F_ID_SQRT, L.F_SINC2_HALF,
                                                             L.F_COSM1_SQRT)

    def __init__(self, row_ptr, col, values, m, complex_vectors=False):
        super().__init__()
        self.rp = np.ascontiguousarray(row_ptr, np.int32)
        self.ci = np.ascontiguousarray(col, np.int32)
        self.val = as_f64(values)
        self.n = self.rp.size - 1
        self.m = m
        self.complex = complex_vectors
        call("nls_matfunc_create", self.n, self.val.size, i32ptr(self.rp), i32ptr(self.ci),
             dptr(self.val), m, int(complex_vectors), C.byref(self.h))

    def last_krylov(self):
        """(T column-major as [col][row], beta0) of the last apply."""
        if self.complex:
            T = np.empty(self.m * self.m, np.complex128)
            buf = T.view(np.float64)
        else:
            T = np.empty(self.m * self.m, np.float64)
            buf = T
        b = C.c_double()
        call("nls_matfunc_last_krylov", self.h, dptr(buf), C.byref(b))
        return T.reshape(self.m, self.m), b.value


class MatrixFunctionApplicatorReal(MatrixFunctionApplicator):
    def __init__(self, row_ptr, col, values, m):
        super().__init__(row_ptr, col, values, m, complex_vectors=False)

    def apply(self, input, t, kind):
        """Host in -> host out: beta0 V f(T) e1 (matfunc_real.hpp:190-313)."""
        from .device import DeviceArray
        x = DeviceArray.from_host(as_f64(input))
        call("nls_matfunc_apply_real", self.h, C.c_void_p(x.ptr), C.c_void_p(x.ptr), float(t),
             int(kind))
        return x.to_host()


class MatrixFunctionApplicatorComplex(MatrixFunctionApplicator):
    EXP, SINC = L.F_EXP, L.F_SINC

    def __init__(self, row_ptr, col, values, m):
        super().__init__(row_ptr, col, values, m, complex_vectors=True)

    def apply(self, input, dt, func="exp"):
        """func "exp": exp(dt L) x ; "sinc": sin(dt L)/(dt L) x (matfunc_complex.hpp:254-375)."""
        from .device import DeviceArray
        kind = {"exp": L.F_EXP, "sinc": L.F_SINC}[func]
        x = DeviceArray.from_host(np.ascontiguousarray(input, np.complex128))
        dt = complex(dt)
        call("nls_matfunc_apply_complex", self.h, C.c_void_p(x.ptr), C.c_void_p(x.ptr), dt.real,
             dt.imag, kind)
        return x.to_host()


def small_matfunc(d, e, kind, t, beta0=1.0):
    """GPU FP64 small stage: y = beta0 Q f(L) Q^T e1 for tridiag(e, d, e)."""
    d = as_f64(d)
    e = as_f64(e) if len(e) else np.zeros(1)
    m = d.size
    yr = np.empty(m)
    yi = np.empty(m)
    t = complex(t)
    call("nls_small_matfunc", m, dptr(d), dptr(e), int(kind), t.real, t.imag, float(beta0),
         dptr(yr), dptr(yi))
    return yr + 1j * yi


class KGESolverDevice(_Handle):
    """Klein-Gordon solver (kg_dev.hpp:19-157).

    method: "gautschi" (step), "sv" (step_sv) or "sewi"; mu2 adds a mass term.
    """
    _destroy = "nls_kge_destroy"
    METHODS = {"gautschi": L.KGE_GAUTSCHI, "sv": L.KGE_SV, "sewi": L.KGE_SEWI}

    def __init__(self, n, L_, T, nt, num_snapshots, m, c, u0, v0, krylov_dim=10,
                 method="gautschi", mu2=0.0, dim=2, q6=True, csr=None, variant="reference"):
        from .batch import lanczos_variant
        super().__init__()
        self.N = n ** dim
        self.shape = (n,) * dim
        self.ns = num_snapshots
        self.p = L.KGEParams(dim, n, L_, T, nt, num_snapshots, krylov_dim,
                             self.METHODS[method], mu2, int(q6), lanczos_variant(variant))
        mf, cf, u0, v0 = (as_f64(a).ravel() for a in (m, c, u0, v0))
        for a in (mf, cf, u0, v0):
            assert a.size == self.N, "field size does not match the grid"
        if csr is None:
            call("nls_kge_create", C.byref(self.p), dptr(mf), dptr(cf), dptr(u0), dptr(v0),
                 C.byref(self.h))
        else:
            rp, ci, val = (np.ascontiguousarray(csr[0], np.int32),
                           np.ascontiguousarray(csr[1], np.int32), as_f64(csr[2]))
            call("nls_kge_create_csr", C.byref(self.p), val.size, i32ptr(rp), i32ptr(ci),
                 dptr(val), dptr(mf), dptr(u0), dptr(v0), C.byref(self.h))

    def step(self):
        call("nls_kge_step", self.h)

    def step_sv(self):
        call("nls_kge_step_sv", self.h)

    def apply_bc(self):
        call("nls_kge_apply_bc", self.h)

    def store_snapshot(self, idx):
        call("nls_kge_store_snapshot", self.h, int(idx))

    def run(self):
        call("nls_kge_run", self.h)

    def transfer_snapshots(self):
        u = np.empty((self.ns,) + self.shape)
        v = np.empty((self.ns,) + self.shape)
        call("nls_kge_transfer_snapshots", self.h, dptr(u), dptr(v))
        return u, v

    def state(self):
        u = np.empty(self.shape)
        v = np.empty(self.shape)
        call("nls_kge_get_state", self.h, dptr(u), dptr(v))
        return u, v


class NLSESolverDevice(_Handle):
    """Cubic NLSE solver (nlse_dev.hpp:66-362); method "strang" or "sewi"."""
    _destroy = "nls_nlse_destroy"

    def __init__(self, n, L_, T, nt, num_snapshots, m, c, u0, krylov_dim=20, method="strang",
                 dim=2, q6=True, variant="reference"):
        from .batch import lanczos_variant
        super().__init__()
        self.N = n ** dim
        self.shape = (n,) * dim
        self.ns = num_snapshots
        meth = {"strang": L.NLSE_STRANG, "sewi": L.NLSE_SEWI}[method]
        self.p = L.NLSEParams(dim, n, L_, T, nt, num_snapshots, krylov_dim, meth, int(q6),
                              lanczos_variant(variant))
        mf, cf = as_f64(m).ravel(), as_f64(c).ravel()
        u = as_c128_view(np.asarray(u0).ravel())
        call("nls_nlse_create", C.byref(self.p), dptr(mf), dptr(cf), dptr(u), C.byref(self.h))

    def step(self, step_number):
        call("nls_nlse_step", self.h, int(step_number))

    def step_sewi(self, step_number):
        call("nls_nlse_step_sewi", self.h, int(step_number))

    def apply_bc(self):
        call("nls_nlse_apply_bc", self.h)

    def run(self):
        call("nls_nlse_run", self.h)

    def transfer_snapshots(self):
        out = np.empty((self.ns,) + self.shape, np.complex128)
        call("nls_nlse_transfer_snapshots", self.h, dptr(out.view(np.float64)))
        return out

    def store_snapshot_online(self, host_dst=None):
        """Stores the current state as the next snapshot (nlse_dev.hpp:324-335); with
        host_dst, a complex128 array of (num_snapshots, *grid), also copies it there."""
        if host_dst is not None:
            assert host_dst.dtype == np.complex128 and host_dst.flags.c_contiguous
            assert host_dst.shape == (self.ns,) + self.shape
        call("nls_nlse_store_snapshot_online", self.h,
             None if host_dst is None else dptr(host_dst.view(np.float64)))

    def state(self):
        out = np.empty(self.shape, np.complex128)
        call("nls_nlse_get_state", self.h, dptr(out.view(np.float64)))
        return out


class SP4SolverDevice(_Handle):
    """Stochastic phi^4 (stochastic_phi4_dev.hpp:17-160) with Philox noise.

    mode "reference": u+ = 2u - u- + tau^2 (L u - m (u - u^3 + s_eff xi)) (the
    reference's form, s_eff = 0.01 sqrt(2 dx)/tau); mode "damped":
    u_tt = L u + m (u - u^3) - gamma u_t + sigma xi.
    """
    _destroy = "nls_sp4_destroy"

    def __init__(self, n, L_, T, nt, num_snapshots, m, c, u0, v0, mode="damped", gamma=0.1,
                 sigma=0.1, seed=0, stream_id=0, dim=2, q6=True):
        super().__init__()
        self.N = n ** dim
        self.shape = (n,) * dim
        self.ns = num_snapshots
        md = {"reference": L.SP4_REFERENCE, "damped": L.SP4_DAMPED}[mode]
        self.p = L.SP4Params(dim, n, L_, T, nt, num_snapshots, md, int(q6), gamma, sigma, seed,
                             stream_id)
        mf, cf, u0, v0 = (as_f64(a).ravel() for a in (m, c, u0, v0))
        call("nls_sp4_create", C.byref(self.p), dptr(mf), dptr(cf), dptr(u0), dptr(v0),
             C.byref(self.h))

    def step(self):
        call("nls_sp4_step", self.h)

    def apply_bc(self):
        call("nls_sp4_apply_bc", self.h)

    def store_snapshot(self, idx):
        call("nls_sp4_store_snapshot", self.h, int(idx))

    def run(self):
        call("nls_sp4_run", self.h)

    def transfer_snapshots(self):
        u = np.empty((self.ns,) + self.shape)
        v = np.empty((self.ns,) + self.shape)
        call("nls_sp4_transfer_snapshots", self.h, dptr(u), dptr(v))
        return u, v


def philox_normals(seed, stream_id, step, count):
    out = np.empty(count)
    call("nls_philox_normals", int(seed), int(stream_id), int(step), int(count), dptr(out))
    return out
