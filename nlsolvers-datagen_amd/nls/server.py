"""TrajectoryServer: continuous batching of stochastic phi^4 requests (nls_server_*)."""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import as_f64, call, dptr


class TrajectoryServer:
    STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}

    def __init__(self, hbm_budget_bytes=8 << 30, max_active=256, L_=10.0, mode="damped",
                 gamma=0.1, max_steps_per_group=64, h5_dir=None, run_id=None):
        """h5_dir: stream every request's snapshots to h5_dir/run_<run_id>_<id:04d>.h5
        (launchers' per-run layout) from a background writer thread."""
        md = {"reference": L.SP4_REFERENCE, "damped": L.SP4_DAMPED}[mode]
        self.p = L.ServerParams(int(hbm_budget_bytes), max_active, L_, md, gamma,
                                max_steps_per_group,
                                None if h5_dir is None else str(h5_dir).encode(),
                                None if run_id is None else str(run_id).encode())
        self.h = C.c_void_p()
        self.ns = {}
        call("nls_server_create", C.byref(self.p), C.byref(self.h))

    def submit(self, n, nt, dt, sigma, seed, num_snapshots=0, m=None, c=None, u0=None, v0=None):
        q = L.Request(n, nt, dt, sigma, seed, num_snapshots)
        arrs = [None if a is None else as_f64(a).ravel() for a in (m, c, u0, v0)]
        ptrs = [None if a is None else dptr(a) for a in arrs]
        rid = C.c_int64()
        call("nls_server_submit", self.h, C.byref(q), *ptrs, C.byref(rid))
        self.ns[rid.value] = (n, num_snapshots)
        return rid.value

    def status(self, rid):
        st = C.c_int()
        call("nls_server_status", self.h, rid, C.byref(st))
        return self.STATES[st.value]

    def wait(self, rid, timeout=600.0):
        call("nls_server_wait", self.h, rid, timeout)

    def drain(self, timeout=600.0):
        call("nls_server_drain", self.h, timeout)

    def fetch(self, rid):
        n, ns = self.ns[rid]
        snaps = np.empty((max(ns, 1), n, n))
        uf = np.empty((n, n))
        cnt = C.c_int()
        lat = C.c_double()
        call("nls_server_fetch", self.h, rid, dptr(snaps), C.byref(cnt), dptr(uf), C.byref(lat))
        return {"snapshots": snaps[:cnt.value], "u_final": uf, "latency_ms": lat.value}

    def release(self, rid):
        call("nls_server_release", self.h, rid)
        self.ns.pop(rid, None)

    def stats(self):
        st = L.ServerStats()
        call("nls_server_get_stats", self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in L.ServerStats._fields_}

    def close(self):
        if self.h and self.h.value:
            call("nls_server_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pool_simulate(capacity, ops):
    """Host-only run of the admission allocator (see nls_pool_simulate)."""
    flat = np.array([x for op in ops for x in op], np.int64)
    res = np.empty(len(ops), np.int64)
    call("nls_pool_simulate", int(capacity), len(ops), flat.ctypes.data_as(L._I64),
         res.ctypes.data_as(L._I64))
    return res.tolist()
