"""CPU checks of the C ABI: libnls.so loads, exports every symbol include/nls.h
declares, and the host-side operator assembly is bit-exact against the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nls.h")).read()
    return sorted(set(re.findall(r"NLS_API\s+[\w\s\*]+?\b(nls_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from nls import _lib
    lib = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 50
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert sorted(_lib.exported_symbols()) == decl


def test_version_and_error_channel():
    from nls import _lib
    lib = _lib.load()
    assert lib.nls_version() >= 100
    # an invalid call fails with a message, without a GPU
    rc = lib.nls_build_laplacian_2d(1, 1.0, 1.0, None, 0, 0.0, None, None, None, None)
    assert rc != 0
    assert b"n must be" in lib.nls_last_error()


def test_operator_assembly_bit_exact_vs_oracle(orc):
    from nls.operators import (build_anisotropic_laplacian_noflux,
                               build_anisotropic_laplacian_noflux_3d)
    rng = np.random.default_rng(0)
    for n in (3, 5, 17, 64):
        dx = 20.0 / (n - 1)
        c = 0.5 + rng.random(n * n)
        for neg, sh in ((False, 0.0), (True, 1.0)):
            rp, ci, v = build_anisotropic_laplacian_noflux(n - 2, n - 2, dx, dx, c, neg, sh)
            op = orc.laplacian_2d(n, dx, dx, c, negate=neg, shift=sh)
            assert np.array_equal(rp, op.rp) and np.array_equal(ci, op.ci)
            assert np.array_equal(v, op.v)
    for n in (3, 4, 9):
        dx = 16.0 / (n - 1)
        c = 0.5 + rng.random(n ** 3)
        for q6 in (True, False):
            rp, ci, v = build_anisotropic_laplacian_noflux_3d(n - 2, n - 2, n - 2, dx, dx, dx, c,
                                                              True, 0.5, q6)
            op = orc.laplacian_3d(n, dx, c, negate=True, shift=0.5, q6=q6)
            assert np.array_equal(rp, op.rp) and np.array_equal(ci, op.ci)
            assert np.array_equal(v, op.v)


def test_nnz_formulas():
    from nls.operators import laplacian_nnz
    for n in (3, 128, 300, 384, 512):
        assert laplacian_nnz(2, n) == n * n + 4 * n * (n - 1)
    assert laplacian_nnz(2, 300) == 448800
    n = 220
    assert laplacian_nnz(3, n, True) == n ** 3 + 2 * (3 * n ** 3 - 2 * n * n - n)
    assert laplacian_nnz(3, n, True) - laplacian_nnz(3, n, False) == 2 * n * (n - 1)


def test_batch_estimate_is_host_only():
    from nls.batch import FusedBatch
    b300 = FusedBatch.estimate_bytes([300] * 100, 10.0, 0.005, krylov_dim=10)
    b150 = FusedBatch.estimate_bytes([300] * 50, 10.0, 0.005, krylov_dim=10)
    assert 1.9 < b300 / b150 < 2.1
    # FP64 state ring + FP32 fields + 2 x (m-2) basis + 2 x 2 w' = 24 + 8 + 64 + 16 bytes/cell
    assert abs(b300 / (100 * 90000) - 112) < 2


def test_batch_parameter_checks_fail_before_any_device_call():
    """Invalid combinations are rejected by nls_batch_create's checks, which run before
    the library touches the device (so this runs without a GPU)."""
    from nls._lib import NLSError
    from nls.batch import FusedBatch
    bad = [dict(reorth=3), dict(reorth="always", pass_kernel=3), dict(pass_kernel=5),
           dict(lanes=5), dict(cells_per_thread=2), dict(always_load_fields=2),
           dict(method="sv", reorth="always"), dict(method="sv", pass_kernel=2)]
    for kw in bad:
        with pytest.raises(NLSError):
            FusedBatch([32], 10.0, 0.005, krylov_dim=8, **kw)
