"""ctypes front-end of the FP64 CPU oracle (oracle/nls_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
Every function here forwards to the C restatement, whose header lists the
reference file:line each routine follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libnls_oracle.so")
_lib = None

F_COS_SQRT, F_SINC2_SQRT, F_ID_SQRT, F_SINC2_HALF, F_COSM1_SQRT = 0, 1, 2, 3, 4

_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_pc = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_pi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class CSR(C.Structure):
    _fields_ = [("N", C.c_int64), ("rp", C.c_void_p), ("ci", C.c_void_p), ("v", C.c_void_p)]


class KGECfg(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("ns", C.c_int), ("krylov", C.c_int), ("method", C.c_int),
                ("variant", C.c_int), ("q6", C.c_int), ("mu2", C.c_double)]


class NLSECfg(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("ns", C.c_int), ("krylov", C.c_int), ("method", C.c_int),
                ("variant", C.c_int), ("q6", C.c_int)]


class SP4Cfg(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("ns", C.c_int), ("mode", C.c_int), ("q6", C.c_int),
                ("gamma", C.c_double), ("sigma", C.c_double), ("seed", C.c_uint64),
                ("stream", C.c_uint32)]


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.orc_lap2d_nnz.restype = C.c_int64
        L.orc_lap2d_nnz.argtypes = [C.c_int]
        L.orc_lap3d_nnz.restype = C.c_int64
        L.orc_lap3d_nnz.argtypes = [C.c_int, C.c_int]
        L.orc_build_lap2d.argtypes = [C.c_int, C.c_double, C.c_double, _pd, C.c_int, C.c_double,
                                      _pi, _pi, _pd]
        L.orc_build_lap3d.argtypes = [C.c_int, C.c_double, _pd, C.c_int, C.c_double, C.c_int,
                                      _pi, _pi, _pd]
        L.orc_csr_spmv.argtypes = [C.c_int64, _pi, _pi, _pd, _pd, _pd]
        L.orc_csr_spmv_c.argtypes = [C.c_int64, _pi, _pi, _pd, _pc, _pc]
        L.orc_tridiag_eig.argtypes = [C.c_int, _pd, _pd, _pd]
        L.orc_lanczos_real.restype = C.c_double
        L.orc_lanczos_real.argtypes = [C.POINTER(CSR), _pd, C.c_int, C.c_int, _pd, _pd]
        L.orc_small_matfunc_real.argtypes = [C.c_int, _pd, C.c_int, C.c_double, _pd]
        L.orc_matfunc_real.argtypes = [C.POINTER(CSR), _pd, _pd, C.c_double, C.c_int, C.c_int,
                                       C.c_int]
        L.orc_lanczos_complex.restype = C.c_double
        L.orc_lanczos_complex.argtypes = [C.POINTER(CSR), _pc, C.c_int, C.c_int, _pc, _pc]
        L.orc_small_matfunc_complex.argtypes = [C.c_int, _pc, C.c_int, C.c_double, C.c_double,
                                                _pc]
        L.orc_matfunc_complex.argtypes = [C.POINTER(CSR), _pc, _pc, C.c_double, C.c_double,
                                          C.c_int, C.c_int, C.c_int]
        L.orc_bc2d.argtypes = [_pd, C.c_int]
        L.orc_bc3d.argtypes = [_pd, C.c_int]
        L.orc_bc2d_c.argtypes = [_pc, C.c_int]
        L.orc_bc3d_c.argtypes = [_pc, C.c_int]
        L.orc_kge_run.argtypes = [C.POINTER(KGECfg), _pd, _pd, _pd, _pd, _pd, _pd]
        L.orc_kge_run_batch.argtypes = [C.c_int, C.POINTER(KGECfg), _pd, _pd, _pd, _pd, _pd, _pd]
        L.orc_nlse_run.argtypes = [C.POINTER(NLSECfg), _pd, _pd, _pc, _pc]
        L.orc_sp4_run.argtypes = [C.POINTER(SP4Cfg), _pd, _pd, _pd, _pd, _pd, _pd]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_det_log.restype = C.c_double
        L.orc_det_log.argtypes = [C.c_double]
        L.orc_det_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_noise_field.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, _pd]
        L.orc_omp_threads.restype = C.c_int
        _lib = L
    return _lib


# --------------------------------------------------------------------------- operator
class Operator:
    """CSR operator kept alive together with the ctypes struct pointing into it."""

    def __init__(self, rp, ci, v):
        self.rp = np.ascontiguousarray(rp, np.int32)
        self.ci = np.ascontiguousarray(ci, np.int32)
        self.v = np.ascontiguousarray(v, np.float64)
        self.N = self.rp.size - 1
        self.s = CSR(self.N, self.rp.ctypes.data, self.ci.ctypes.data, self.v.ctypes.data)

    def dense(self):
        A = np.zeros((self.N, self.N))
        for r in range(self.N):
            for k in range(self.rp[r], self.rp[r + 1]):
                A[r, self.ci[k]] = self.v[k]
        return A

    def matvec(self, x):
        x = np.ascontiguousarray(x)
        if np.iscomplexobj(x):
            y = np.empty(self.N, np.complex128)
            lib().orc_csr_spmv_c(self.N, self.rp, self.ci, self.v, x.astype(np.complex128), y)
        else:
            y = np.empty(self.N, np.float64)
            lib().orc_csr_spmv(self.N, self.rp, self.ci, self.v, x.astype(np.float64), y)
        return y


def laplacian_2d(n, dx, dy, c, negate=False, shift=0.0):
    """Full-array 2D operator (laplacians.hpp:54-103), n = full side."""
    c = np.ascontiguousarray(c, np.float64).ravel()
    nnz = lib().orc_lap2d_nnz(n)
    rp = np.empty(n * n + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    lib().orc_build_lap2d(n, dx, dy, c, int(negate), float(shift), rp, ci, v)
    return Operator(rp, ci, v)


def laplacian_3d(n, dx, c, negate=False, shift=0.0, q6=True):
    """Full-array 3D operator (laplacians.hpp:158-218), Q6 cross-plane couplings if q6."""
    c = np.ascontiguousarray(c, np.float64).ravel()
    nnz = lib().orc_lap3d_nnz(n, int(q6))
    rp = np.empty(n ** 3 + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    lib().orc_build_lap3d(n, dx, c, int(negate), float(shift), int(q6), rp, ci, v)
    return Operator(rp, ci, v)


# --------------------------------------------------------------------------- Krylov
def tridiag_eig(d, e):
    d = np.array(d, np.float64)
    e = np.array(e, np.float64)
    m = d.size
    Z = np.empty(m * m, np.float64)
    rc = lib().orc_tridiag_eig(m, d, np.ascontiguousarray(np.append(e, 0.0)), Z)
    if rc:
        raise RuntimeError("tridiagonal QL did not converge")
    return d, Z.reshape(m, m)


def lanczos_real(op, x, m, variant=0):
    V = np.empty(op.N * m, np.float64)
    T = np.empty(m * m, np.float64)
    beta = lib().orc_lanczos_real(C.byref(op.s), np.ascontiguousarray(x, np.float64), m,
                                  variant, V, T)
    return beta, V.reshape(m, op.N), T.reshape(m, m)  # T[col, row] (column-major)


def small_matfunc_real(Tcm, kind, t):
    m = Tcm.shape[0]
    y = np.empty(m, np.float64)
    rc = lib().orc_small_matfunc_real(m, np.ascontiguousarray(Tcm.ravel()), kind, t, y)
    if rc:
        raise RuntimeError("eigensolver failed")
    return y


def matfunc_real(op, x, t, kind, m, variant=0):
    out = np.empty(op.N, np.float64)
    rc = lib().orc_matfunc_real(C.byref(op.s), out, np.ascontiguousarray(x, np.float64), t, kind,
                                m, variant)
    if rc:
        raise RuntimeError("matfunc failed")
    return out


def matfunc_complex(op, x, dt, kind, m, variant=0):
    out = np.empty(op.N, np.complex128)
    dt = complex(dt)
    rc = lib().orc_matfunc_complex(C.byref(op.s), out, np.ascontiguousarray(x, np.complex128),
                                   dt.real, dt.imag, kind, m, variant)
    if rc:
        raise RuntimeError("matfunc failed")
    return out


def small_matfunc_complex(Tcm, kind, dt):
    m = Tcm.shape[0]
    y = np.empty(m, np.complex128)
    dt = complex(dt)
    rc = lib().orc_small_matfunc_complex(m, np.ascontiguousarray(Tcm.ravel(), np.complex128),
                                         kind, dt.real, dt.imag, y)
    if rc:
        raise RuntimeError("eigensolver failed")
    return y


def bc2d(u):
    u = np.ascontiguousarray(u).copy()
    n = u.shape[0]
    if np.iscomplexobj(u):
        lib().orc_bc2d_c(u.ravel(), n)
    else:
        lib().orc_bc2d(u.ravel(), n)
    return u


def bc3d(u):
    u = np.ascontiguousarray(u).copy()
    n = u.shape[0]
    if np.iscomplexobj(u):
        lib().orc_bc3d_c(u.ravel(), n)
    else:
        lib().orc_bc3d(u.ravel(), n)
    return u


# --------------------------------------------------------------------------- drivers
def kge_run(n, L, T, nt, ns, m_field, c, u0, v0, krylov=10, method=0, mu2=0.0, variant=0,
            dim=2, q6=True):
    N = n ** dim
    cfg = KGECfg(dim, n, L, T, nt, ns, krylov, method, variant, int(q6), mu2)
    ou = np.zeros(ns * N, np.float64)
    ov = np.zeros(ns * N, np.float64)
    rc = lib().orc_kge_run(C.byref(cfg), _flat(m_field), _flat(c), _flat(u0), _flat(v0), ou, ov)
    if rc:
        raise RuntimeError("oracle kge_run failed")
    shape = (ns,) + (n,) * dim
    return ou.reshape(shape), ov.reshape(shape)


def kge_run_batch(n, L, T, nt, ns, m_fields, cs, u0s, v0s, krylov=10, method=0, mu2=0.0,
                  variant=0, dim=2, q6=True):
    B = len(u0s)
    N = n ** dim
    cfg = KGECfg(dim, n, L, T, nt, ns, krylov, method, variant, int(q6), mu2)
    ou = np.zeros(B * ns * N, np.float64)
    ov = np.zeros(B * ns * N, np.float64)
    rc = lib().orc_kge_run_batch(B, C.byref(cfg), _stack(m_fields), _stack(cs), _stack(u0s),
                                 _stack(v0s), ou, ov)
    if rc:
        raise RuntimeError("oracle kge_run_batch failed")
    shape = (B, ns) + (n,) * dim
    return ou.reshape(shape), ov.reshape(shape)


def nlse_run(n, L, T, nt, ns, m_field, c, u0, krylov=12, method=0, variant=0, dim=2, q6=True):
    N = n ** dim
    cfg = NLSECfg(dim, n, L, T, nt, ns, krylov, method, variant, int(q6))
    out = np.zeros(ns * N, np.complex128)
    rc = lib().orc_nlse_run(C.byref(cfg), _flat(m_field), _flat(c),
                            np.ascontiguousarray(u0, np.complex128).ravel(), out)
    if rc:
        raise RuntimeError("oracle nlse_run failed")
    return out.reshape((ns,) + (n,) * dim)


def sp4_run(n, L, T, nt, ns, m_field, c, u0, v0, mode=1, gamma=0.1, sigma=0.1, seed=0,
            stream=0, dim=2, q6=True):
    N = n ** dim
    cfg = SP4Cfg(dim, n, L, T, nt, ns, mode, int(q6), gamma, sigma, seed, stream)
    ou = np.zeros(ns * N, np.float64)
    ov = np.zeros(ns * N, np.float64)
    lib().orc_sp4_run(C.byref(cfg), _flat(m_field), _flat(c), _flat(u0), _flat(v0), ou, ov)
    shape = (ns,) + (n,) * dim
    return ou.reshape(shape), ov.reshape(shape)


# --------------------------------------------------------------------------- noise
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def det_log(u):
    return lib().orc_det_log(float(u))


def det_sincos2pi(w):
    cs, sn = C.c_double(), C.c_double()
    lib().orc_det_sincos2pi(float(w), C.byref(cs), C.byref(sn))
    return cs.value, sn.value


def noise_field(seed, stream, step, N):
    z = np.empty(N, np.float64)
    lib().orc_noise_field(seed, stream, step, N, z)
    return z


def omp_threads():
    return lib().orc_omp_threads()


def _flat(a):
    return np.ascontiguousarray(a, np.float64).ravel()


def _stack(arrs):
    return np.ascontiguousarray(np.stack([np.asarray(a, np.float64).ravel() for a in arrs]))
