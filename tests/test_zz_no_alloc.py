"""The allocation-free matrix-function entry point (nls_matfunc_apply_real_noalloc)
against the handle-based one (nls_matfunc_apply_real) on the same operator and input:
the same kernels on a view of the fixed device scratch, so the results are bitwise equal;
two calls in a row, the second on a smaller operator, and the oversize rejection."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def last_error():
    from nls import _lib as L
    return L.load().nls_last_error().decode(errors="replace")


def _lib():
    from nls import _lib as L
    return L, L.load().nls_matfunc_apply_real_noalloc


def _case(orc, n, seed):
    from nls.samplers import gaussian_sum_2d
    rng = np.random.default_rng(seed)
    c = 0.8 + 0.4 * rng.random((n, n))
    L = 10.0
    dx = 2 * L / (n - 1)
    op = orc.laplacian_2d(n, dx, dx, c.ravel(), negate=True, shift=1.0)
    u0, _, _ = gaussian_sum_2d(n, L, seed, 0)
    return op, u0.ravel().copy()


def _handle_result(L, op, x, t, kind, m):
    from nls._lib import call, dptr
    N = len(x)
    h = C.c_void_p()
    rp, ci, v = (np.ascontiguousarray(a) for a in (op.rp, op.ci, op.v))
    call("nls_matfunc_create", N, len(ci), rp.ctypes.data_as(C.POINTER(C.c_int32)),
         ci.ctypes.data_as(C.POINTER(C.c_int32)), dptr(v), m, 0, C.byref(h))
    from nls.device import DeviceArray
    dx_, dy_ = DeviceArray.from_host(x), DeviceArray(N, np.float64)
    call("nls_matfunc_apply_real", h, dy_.ptr, dx_.ptr, t, kind)
    out = dy_.to_host()
    call("nls_matfunc_destroy", h)
    return out


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_matfunc_equals_handle_entry(orc, kind):
    from nls.device import DeviceArray
    L, f = _lib()
    t, m = 0.01, 10
    for n, seed in ((128, 1), (48, 2)):  # the second call on a smaller operator
        op, x = _case(orc, n, seed)
        N = n * n
        rp = DeviceArray.from_host(np.asarray(op.rp, dtype=np.int32))
        ci = DeviceArray.from_host(np.asarray(op.ci, dtype=np.int32))
        val = DeviceArray.from_host(np.asarray(op.v, dtype=np.float64))
        dx_, dy_ = DeviceArray.from_host(x), DeviceArray(N, np.float64)
        rc = f(N, rp.ptr, ci.ptr, val.ptr, dy_.ptr, dx_.ptr, t, kind, m, None)
        assert rc == 0, last_error()
        got = dy_.to_host()
        ref = _handle_result(L, op, x, t, kind, m)
        assert np.array_equal(got, ref), (n, np.abs(got - ref).max())
        want = orc.matfunc_real(op, x, t, kind, m)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), n


def test_noalloc_matfunc_rejects_oversize():
    L, f = _lib()
    n, m = 200_000, 32  # 8 n (m + 1) bytes > 32 MiB; rejected before any device access
    fake = 256
    rc = f(n, fake, fake, fake, fake, fake, 0.01, 0, m, None)
    assert rc != 0
    assert "scratch" in last_error()
