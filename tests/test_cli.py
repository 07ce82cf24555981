"""Reference-named command-line drivers (device/drivers/*.cpp argv): CPU error
paths, and GPU runs against the oracle through .npy files."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "nlsolvers-datagen_amd", "bin")


def exe(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} missing: run __graft_entry__.build()")
    return p


def test_cli_usage_and_shape_errors(tmp_path):
    n = 8
    for f, shp in (("u0", (n, n)), ("m", (n, n)), ("c", (n, n)), ("bad", (n, n + 1))):
        np.save(tmp_path / f"{f}.npy", np.ones(shp))
    r = subprocess.run([exe("kg_gautschi_2d_dev"), "8", "8"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    t = str(tmp_path)
    r = subprocess.run([exe("kg_gautschi_2d_dev"), "8", "8", "5", "5", f"{t}/bad.npy", f"{t}/u0.npy",
                        f"{t}/tr.npy", f"{t}/ve.npy", "1", "10", "2", f"{t}/m.npy", f"{t}/c.npy"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "mismatch" in r.stderr
    # a wrong-shape m throws, as the reference driver does (kg_driver_dev_2d.cpp:75-84)
    r = subprocess.run([exe("kg_gautschi_2d_dev"), "8", "8", "5", "5", f"{t}/u0.npy", f"{t}/u0.npy",
                        f"{t}/tr.npy", f"{t}/ve.npy", "1", "10", "2", f"{t}/bad.npy", f"{t}/c.npy"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "Faulty m" in r.stderr


@pytest.mark.gpu
def test_cli_kg_and_sv_vs_oracle(orc, tmp_path):
    from nls.samplers import gaussian_sum_2d
    n, L, T, nt, ns = 48, 10.0, 0.5, 50, 5
    u0, v0, _ = gaussian_sum_2d(n, L, 1, 0)
    rng = np.random.default_rng(0)
    m = 0.5 + rng.random((n, n))
    c = 0.5 + rng.random((n, n))
    t = str(tmp_path)
    for name, arr in (("u0", u0), ("v0", v0), ("m", m), ("c", c)):
        np.save(f"{t}/{name}.npy", arr)
    for exe_name, method in (("kg_gautschi_2d_dev", 0), ("kg_sv_2d_dev", 1)):
        r = subprocess.run([exe(exe_name), str(n), str(n), str(L), str(L), f"{t}/u0.npy",
                            f"{t}/v0.npy", f"{t}/traj.npy", f"{t}/vel.npy", str(T), str(nt),
                            str(ns), f"{t}/m.npy", f"{t}/c.npy"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        u = np.load(f"{t}/traj.npy")
        v = np.load(f"{t}/vel.npy")
        assert u.shape == (ns, n, n) and u.dtype == np.float64
        ur, vr = orc.kge_run(n, L, T, nt, ns, m, c, u0, v0, krylov=10, method=method)
        assert np.linalg.norm(u - ur) / np.linalg.norm(ur) < 1e-10
        assert np.linalg.norm(v[1:] - vr[1:]) / np.linalg.norm(vr[1:]) < 1e-8


@pytest.mark.gpu
def test_cli_nlse_vs_oracle(orc, tmp_path):
    from nls.samplers import nlse_packets_2d
    n, L, T, nt, ns = 40, 10.0, 0.2, 40, 4
    psi, _ = nlse_packets_2d(n, L, 1, 3)
    t = str(tmp_path)
    np.save(f"{t}/u0.npy", psi)
    np.save(f"{t}/m.npy", np.ones((n, n)))
    np.save(f"{t}/c.npy", np.ones((n, n)))
    r = subprocess.run([exe("nlse_2d_dev"), str(n), str(n), str(L), str(L), f"{t}/u0.npy",
                        f"{t}/traj.npy", str(T), str(nt), str(ns), f"{t}/m.npy", f"{t}/c.npy"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = np.load(f"{t}/traj.npy")
    assert out.shape == (ns, n, n) and out.dtype == np.complex128
    ref = orc.nlse_run(n, L, T, nt, ns, np.ones((n, n)), np.ones((n, n)), psi, krylov=20,
                       method=0, variant=0)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-10
