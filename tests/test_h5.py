"""HDF5 per-run files in the launchers' layout (scripts/real_launcher_2d.py:201-245,
scripts/real_launcher_3d.py:206-252, scripts/complex_launcher_2d.py:193-236): round
trip through libnls' reader, the stored types checked with h5dump (the HDF5 tool),
and the asynchronous writer.  Host-only: no GPU is used."""
import os
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def h5():
    from nls import h5 as mod
    if mod.available() is None:
        pytest.skip("libhdf5 not loadable here")
    return mod


def real_case(n=12, ns=3, seed=0):
    rng = np.random.default_rng(seed)
    return dict(u0=rng.standard_normal((n, n)), v0=rng.standard_normal((n, n)),
                m=rng.random((n, n)), c=1 + rng.random((n, n)),
                u=rng.standard_normal((ns, n, n)), v=rng.standard_normal((ns, n, n)))


def test_real_run_round_trip(h5, tmp_path):
    d = real_case()
    p = str(tmp_path / "run_abcd1234_0003.h5")
    h5.save_run(p, **d, Lx=10.0, T=5.0, nt=500, num_snapshots=3, problem_type="klein_gordon",
                run_id="abcd1234", run_index=3, phenomenon="kink_field",
                phenomenon_params={"radius": 2.5, "centers": [(0, 1), (2, 3)]},
                m_type="constant", elapsed_time=1.25, timestamp="2025-01-02 03:04:05.000006")
    for name, key in (("initial_condition/u0", "u0"), ("initial_condition/v0", "v0"),
                      ("focusing/m", "m"), ("focusing/c", "c"), ("u", "u"), ("v", "v")):
        assert np.array_equal(h5.read(p, name), d[key]), name
    x = np.linspace(-10, 10, 12)
    X, Y = np.meshgrid(x, x, indexing="ij")
    assert np.array_equal(h5.read(p, "X"), X) and np.array_equal(h5.read(p, "Y"), Y)
    assert h5.read_attr(p, "metadata", "problem_type") == "klein_gordon"
    assert h5.read_attr(p, "metadata", "boundary_condition") == "noflux"
    assert h5.read_attr(p, "metadata", "run_id") == "abcd1234"
    assert h5.read_attr(p, "metadata", "run_index") == 3
    assert h5.read_attr(p, "metadata", "elapsed_time") == 1.25
    assert h5.read_attr(p, "metadata", "timestamp") == "2025-01-02 03:04:05.000006"
    assert h5.read_attr(p, "metadata", "phenomenon") == "kink_field"
    assert h5.read_attr(p, "metadata", "phenomenon_radius") == "2.5"
    assert h5.read_attr(p, "metadata", "phenomenon_centers") == str([(0, 1), (2, 3)])
    assert h5.read_attr(p, "grid", "nx") == 12 and h5.read_attr(p, "grid", "Ly") == 10.0
    assert h5.read_attr(p, "time", "T") == 5.0 and h5.read_attr(p, "time", "nt") == 500
    assert h5.read_attr(p, "time", "num_snapshots") == 3
    assert h5.read_attr(p, "focusing", "type") == "constant"


def test_complex_run_layout(h5, tmp_path):
    rng = np.random.default_rng(1)
    n, ns = 10, 2
    u0 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    u = rng.standard_normal((ns, n, n)) + 1j * rng.standard_normal((ns, n, n))
    m, c = rng.random((n, n)), rng.random((n, n))
    p = str(tmp_path / "c.h5")
    h5.save_run(p, u0=u0, m=m, c=c, u=u, Lx=8.0, T=1.0, nt=100, num_snapshots=2,
                problem_type="cubic", run_id="r", run_index=0, phenomenon="multi_soliton")
    assert np.array_equal(h5.read(p, "initial_condition/u0"), u0)
    assert np.array_equal(h5.read(p, "u"), u)
    assert np.array_equal(h5.read(p, "focusing/m"), m)
    assert np.array_equal(h5.read(p, "c"), c)  # complex launcher: c at the root
    with pytest.raises(Exception):
        h5.read(p, "v")


def test_3d_run_has_3d_meshgrid(h5, tmp_path):
    rng = np.random.default_rng(2)
    n = 6
    g = lambda: rng.standard_normal((n, n, n))
    p = str(tmp_path / "r3.h5")
    h5.save_run(p, u0=g(), v0=g(), m=g(), c=g(), u=rng.standard_normal((2, 4, 4, 4)),
                v=rng.standard_normal((2, 4, 4, 4)), Lx=8.0, Ly=8.0, Lz=8.0, T=1.0, nt=10,
                num_snapshots=2, problem_type="klein_gordon", run_id="r", run_index=1)
    x = np.linspace(-8, 8, n)
    X, Y, _ = np.meshgrid(x, x, x, indexing="ij")
    assert np.array_equal(h5.read(p, "X"), X) and np.array_equal(h5.read(p, "Y"), Y)
    assert h5.read(p, "u").shape == (2, 4, 4, 4)
    assert h5.read_attr(p, "grid", "nz") == 6 and h5.read_attr(p, "grid", "Lz") == 8.0


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_types_match_h5py_conventions(h5, tmp_path):
    """Strings variable-length UTF-8, ints 64-bit, complex a compound {r, i} of
    doubles: the types h5py writes for str, int and complex128."""
    d = real_case(n=5, ns=1)
    p = str(tmp_path / "t.h5")
    h5.save_run(p, **d, Lx=1.0, T=1.0, nt=2, num_snapshots=1, problem_type="phi4", run_id="x",
                run_index=7)
    out = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, check=True).stdout
    assert 'ATTRIBUTE "problem_type"' in out and "H5T_VARIABLE" in out and "H5T_CSET_UTF8" in out
    assert "H5T_STD_I64LE" in out and "H5T_IEEE_F64LE" in out
    for grp in ("metadata", "grid", "time", "initial_condition", "focusing"):
        assert f'GROUP "{grp}"' in out
    for ds in ('DATASET "u"', 'DATASET "v"', 'DATASET "X"', 'DATASET "Y"', 'DATASET "u0"',
               'DATASET "v0"', 'DATASET "m"', 'DATASET "c"'):
        assert ds in out
    pc = str(tmp_path / "tc.h5")
    z = np.ones((4, 4), np.complex128)
    h5.save_run(pc, u0=z, m=np.ones((4, 4)), c=np.ones((4, 4)), u=z[None], Lx=1.0, T=1.0, nt=2,
                num_snapshots=1, problem_type="cubic", run_id="x", run_index=0)
    outc = subprocess.run([H5DUMP, "-H", pc], capture_output=True, text=True, check=True).stdout
    assert "H5T_COMPOUND" in outc and '"r"' in outc and '"i"' in outc


def test_async_writer_writes_every_file(h5, tmp_path):
    cases = [real_case(n=16, ns=4, seed=s) for s in range(6)]
    with h5.AsyncWriter(max_queued_bytes=1 << 16) as w:  # small queue: submit waits
        for k, d in enumerate(cases):
            w.submit(str(tmp_path / f"run_q_{k:04d}.h5"), **d, Lx=10.0, T=1.0, nt=40,
                     num_snapshots=4, problem_type="klein_gordon", run_id="q", run_index=k)
        assert w.drain() == len(cases)
    for k, d in enumerate(cases):
        assert np.array_equal(h5.read(str(tmp_path / f"run_q_{k:04d}.h5"), "u"), d["u"])


def test_writer_reports_failures(h5, tmp_path):
    d = real_case(n=4, ns=1)
    w = h5.AsyncWriter()
    w.submit(str(tmp_path / "missing_dir" / "x.h5"), **d, Lx=1.0, T=1.0, nt=2, num_snapshots=1,
             problem_type="p", run_id="r", run_index=0)
    with pytest.raises(Exception, match="HDF5"):
        w.drain()
    with pytest.raises(Exception, match="HDF5"):
        w.close()  # drains again (same failure), then releases the writer
    assert not w.h


def test_bad_shapes_are_rejected(h5, tmp_path):
    d = real_case(n=6)
    d["m"] = np.ones((5, 5))
    with pytest.raises(ValueError):
        h5.save_run(str(tmp_path / "bad.h5"), **d, Lx=1.0, T=1.0, nt=2, num_snapshots=1,
                    problem_type="p", run_id="r", run_index=0)
