"""FusedBatch: many trajectories advanced together by the fused kernels.

Replaces running one solver process per trajectory (the reference's
scripts/real_launcher_2d.py:260-306 loop over subprocess calls).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import as_f64, call, dptr


IC_KINDS = {"gaussians_2d": 0, "kink_ring_2d": 1, "packets_2d": 2, "bubbles_3d": 3}


def lanczos_variant(v):
    """Lanczos variant code: "reference" (0) keeps the reference's T[m-1][m-1] = 0
    (device/include/lanczos.hpp:234-263); "full" (1) computes alpha_{m-1} with one more
    pass per step, the full tridiagonal T."""
    return {"reference": 0, "ref": 0, "full": 1, "textbook": 1}.get(v, v)


def ic_params(kind, seed, traj, L=None):
    """The sampler parameters drawn for one trajectory on [-L, L]^dim (default L: 10 in
    2D, 8 in 3D, the configurations' domains), as the library draws them (host only; see
    nls_ic_draw for the layout)."""
    L = (8.0 if kind == "bubbles_3d" else 10.0) if L is None else float(L)
    buf = np.empty(32)
    cnt = C.c_int()
    call("nls_ic_draw", IC_KINDS[kind], int(seed), int(traj), L, dptr(buf), 32, C.byref(cnt))
    return buf[:cnt.value].copy()


class FusedBatch:
    def __init__(self, ns_list, Ls, dt, equation="kge", method="gautschi", dim=2, krylov_dim=0,
                 mu2=0.0, q6=True, gamma=0.0, use_graph=True, snapshot_every=0,
                 max_snapshots=0, sigmas=None, seeds=None, cells_per_thread=0,
                 pass_kernel=None, lanes=0, reorth=0, reorth_tol=0.0,
                 always_load_fields=False, variant="reference", persistent=False):
        eq = {"kge": L.EQ_KGE, "nlse": L.EQ_NLSE, "sp4": L.EQ_SP4}[equation]
        meth = {"gautschi": L.KGE_GAUTSCHI, "sewi": L.KGE_SEWI, "sv": L.KGE_SV,
                "strang": L.NLSE_STRANG, "reference": L.SP4_REFERENCE,
                "damped": L.SP4_DAMPED}[method]
        self.dim = dim
        self.equation = equation
        self.ns = list(ns_list)
        k = len(self.ns)
        Ls = list(Ls) if np.ndim(Ls) else [Ls] * k
        sigmas = list(sigmas) if sigmas is not None else [0.0] * k
        seeds = list(seeds) if seeds is not None else [0] * k
        self.p = L.BatchParams(eq, meth, dim, dt, krylov_dim, mu2, int(q6), gamma, int(use_graph),
                               snapshot_every, max_snapshots, cells_per_thread,
                               0 if pass_kernel is None else pass_kernel + 1, lanes,
                               {"never": 0, "always": 1, "auto": 2}.get(reorth, reorth),
                               reorth_tol, int(always_load_fields), lanczos_variant(variant),
                               int(persistent))
        self.td = (L.TrajDesc * k)(*[L.TrajDesc(n, l, s, sd)
                                     for n, l, s, sd in zip(self.ns, Ls, sigmas, seeds)])
        self.h = C.c_void_p()
        call("nls_batch_create", C.byref(self.p), k, self.td, C.byref(self.h))

    @staticmethod
    def estimate_bytes(ns_list, Ls, dt, dim=2, krylov_dim=0, mu2=0.0, max_snapshots=0,
                       equation="kge", variant="reference"):
        k = len(ns_list)
        Ls = list(Ls) if np.ndim(Ls) else [Ls] * k
        eq = {"kge": L.EQ_KGE, "nlse": L.EQ_NLSE, "sp4": L.EQ_SP4}[equation]
        p = L.BatchParams(eq, 0, dim, dt, krylov_dim, mu2, 1, 0.0, 1, 0, max_snapshots, 0, 0, 0,
                          0, 0.0, 0, lanczos_variant(variant))
        td = (L.TrajDesc * k)(*[L.TrajDesc(n, l, 0.0, 0) for n, l in zip(ns_list, Ls)])
        out = C.c_size_t()
        call("nls_batch_estimate_bytes", C.byref(p), k, td, C.byref(out))
        return out.value

    def set_initial(self, t, m, c, u0, v0=None):
        N = self.ns[t] ** self.dim
        mf, cf = as_f64(m).ravel(), as_f64(c).ravel()
        if self.equation == "nlse":
            u = np.ascontiguousarray(u0, np.complex128).ravel().view(np.float64)
            v = np.zeros(N)
        else:
            u = as_f64(u0).ravel()
            v = as_f64(v0 if v0 is not None else np.zeros(N)).ravel()
        call("nls_batch_set_initial", self.h, t, dptr(mf), dptr(cf), dptr(u), dptr(v))

    def sample_initial(self, t, kind, seed, traj_index=None, m=1.0, c=1.0):
        """Initial data of trajectory t drawn and evaluated on the GPU (no host arrays):
        kind one of IC_KINDS ("gaussians_2d", "kink_ring_2d", "packets_2d", "bubbles_3d",
        the fields of nls.samplers' gaussian_sum_2d, kink_ring_2d, nlse_packets_2d,
        bubbles_3d for the same seed and trajectory counter), v0 = 0, constant m and c.
        traj_index (default t) is the sampler's trajectory counter."""
        call("nls_batch_set_initial_sampled", self.h, t, IC_KINDS[kind], int(seed),
             t if traj_index is None else int(traj_index), float(m), float(c))

    def run(self, nsteps):
        ms = C.c_float()
        call("nls_batch_run", self.h, int(nsteps), C.byref(ms))
        return ms.value

    def snapshots(self, t):
        N = self.ns[t] ** self.dim
        cnt = C.c_int()
        call("nls_batch_get_snapshots", self.h, t, None, None, C.byref(cnt))
        shape = (cnt.value,) + (self.ns[t],) * self.dim
        if self.equation == "nlse":
            u = np.empty(shape, np.complex128)
            call("nls_batch_get_snapshots", self.h, t, dptr(u.view(np.float64)), None, C.byref(cnt))
            return u, None
        u = np.empty(shape)
        v = np.empty(shape)
        call("nls_batch_get_snapshots", self.h, t, dptr(u), dptr(v), C.byref(cnt))
        assert N * cnt.value == u.size
        return u, v

    def state(self, t):
        shape = (self.ns[t],) * self.dim
        if self.equation == "nlse":
            u = np.empty(shape, np.complex128)
            call("nls_batch_get_state", self.h, t, dptr(u.view(np.float64)), None)
            return u, None
        u = np.empty(shape)
        v = np.empty(shape)
        call("nls_batch_get_state", self.h, t, dptr(u), dptr(v))
        return u, v

    def info(self):
        m = C.c_int()
        cells = C.c_int64()
        b = C.c_size_t()
        call("nls_batch_info", self.h, C.byref(m), C.byref(cells), C.byref(b))
        return {"krylov_dim": m.value, "cells": cells.value, "device_bytes": b.value}

    def tile(self):
        cpt = C.c_int()
        kind = C.c_int()
        lanes = C.c_int()
        call("nls_batch_tile", self.h, C.byref(cpt), C.byref(kind), C.byref(lanes))
        return {"cells_per_thread": cpt.value, "pass_kind": kind.value, "lanes": lanes.value}

    def persistent_steps(self):
        """Steps run so far by the persistent step kernel (``persistent=True``)."""
        n = C.c_int64()
        call("nls_batch_persistent_steps", self.h, C.byref(n))
        return n.value

    def abort(self):
        """Ends the persistent launch in progress (or the next one): that run raises
        NLSError; the batch is then left mid-step and should be closed."""
        call("nls_batch_abort", self.h)

    def monitor(self, t):
        """Krylov monitor of trajectory t: steps whose last-vector weight |y_{m-1}|/|y|
        exceeded 1e-6, the largest weight seen, the largest loss of orthogonality
        |<v_{m-2}, v_0>| seen (reorthogonalisation modes 1 and 2) and whether full
        reorthogonalisation is active."""
        st = C.c_int()
        mt = C.c_double()
        mo = C.c_double()
        on = C.c_int()
        call("nls_batch_monitor", self.h, t, C.byref(st), C.byref(mt), C.byref(mo), C.byref(on))
        return {"stalls": st.value, "max_tail": mt.value, "max_orth": mo.value,
                "reorth_on": bool(on.value)}

    def set_profiling(self, on):
        call("nls_batch_set_profiling", self.h, int(bool(on)))

    def pass_timing(self):
        avg = C.c_double()
        cnt = C.c_int()
        byt = C.c_double()
        call("nls_batch_pass_timing", self.h, C.byref(avg), C.byref(cnt), C.byref(byt))
        return {"avg_ms": avg.value, "launches": cnt.value, "bytes_per_launch": byt.value}

    def close(self):
        if self.h and self.h.value:
            call("nls_batch_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
