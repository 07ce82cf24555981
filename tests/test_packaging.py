"""Dataset packaging (host only): YAML descriptors checked key for key against the
reference's own dataset descriptors (tests/golden/reference_configs/, copies of the
reference's config_*.yaml), and per-class NetCDF-4 files checked for the layout of
aggregate_single_core.py:167-206 with h5dump and for their contents by reading them back.
(No netCDF4 reader is importable here: that a netCDF library reads the files is not
pinned.)"""
import os
import shutil
import subprocess

import numpy as np
import pytest
import yaml

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_configs")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def h5():
    from nls import h5 as mod
    if mod.available() is None:
        pytest.skip("libhdf5 not loadable")
    return mod


def _ref(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return yaml.safe_load(f)


def _write_runs(h5, root, n, L, T, nt, ns, classes, problem_type="klein_gordon", out_n=4,
                cplx=False, start=0, nan_at=None):
    """One run file per (m class, c class) pair in `classes`, stored trajectory of
    ns x out_n x out_n; returns the paths."""
    rng = np.random.default_rng(start)
    paths = []
    for k, (mt, ct) in enumerate(classes):
        i = start + k
        u0 = rng.standard_normal((n, n))
        u = rng.standard_normal((ns, out_n, out_n))
        if cplx:
            u0 = u0 + 1j * rng.standard_normal((n, n))
            u = u + 1j * rng.standard_normal((ns, out_n, out_n))
        if nan_at == k:
            u[1, 0, 0] = np.nan
        kw = dict(u0=u0, m=rng.random((n, n)) + 0.5, c=rng.random((n, n)) + 0.5, u=u, Lx=L,
                  T=T, nt=nt, num_snapshots=ns, problem_type=problem_type, run_id="x",
                  run_index=i, phenomenon="p", m_type=mt, c_type=ct)
        if not cplx:
            kw.update(v0=rng.standard_normal((n, n)), v=rng.standard_normal((ns, out_n, out_n)))
        p = os.path.join(root, f"run_x_{i:04d}.h5")
        h5.save_run(p, **kw)
        paths.append(p)
    return paths


def test_descriptor_matches_reference_config_files(h5, tmp_path):
    """Run files with the 2D Klein-Gordon reference dataset's parameters (300^2 on
    [-3, 3]^2, T = 20, nt = 15000, 100 snapshots, its eight m and seven c classes) give
    the reference's descriptor: same file name, same keys in the same order, same values
    (apart from the trajectory count)."""
    from nls.packaging import write_descriptors
    ref_name = "config_Klein-Gordon_2D_300x300_ns100_nt15000.yaml"
    ref = _ref(ref_name)
    ms, cs = ref["m_classes"], ref["c_classes"]
    pairs = [(ms[k % len(ms)], cs[k % len(cs)]) for k in range(max(len(ms), len(cs)))]
    runs = tmp_path / "runs"
    runs.mkdir()
    _write_runs(h5, str(runs), 300, 3.0, 20.0, 15000, 100, pairs)
    out = write_descriptors(str(runs), str(tmp_path / "cfg"))
    assert [os.path.basename(p) for p in out] == [ref_name]
    with open(out[0]) as f:
        got = yaml.safe_load(f)
    assert list(got) == list(ref)
    assert got["num_trajectories"] == len(pairs)
    for k in ref:
        if k != "num_trajectories":
            assert got[k] == ref[k], k


@pytest.mark.parametrize("name", ["config_Nonlinear_Schrodinger_2D_200x200_ns128_nt2000.yaml",
                                  "config_Klein-Gordon_3D_220x220x220_ns40_nt10000.yaml",
                                  "config_Klein-Gordon_2D_300x300_ns50_nt6000.yaml"])
def test_descriptor_function_reproduces_reference_configs(name):
    """The descriptor of run metadata equal to each other reference dataset's (the NLSE
    and 3D ones would need full-size arrays as files): name, keys, order, values."""
    from nls.packaging import descriptors
    ref = _ref(name)
    three = "nz" in ref["grid"]
    metas = []
    ms = ref.get("m_classes", [None])
    cs = ref.get("c_classes", [None])
    for k in range(max(len(ms), len(cs))):
        metas.append({"dims": 3 if three else 2, "pde_type": ref["pde_type"],
                      "boundary": ref["boundary_condition"], "nx": ref["grid"]["nx"],
                      "ny": ref["grid"]["ny"], "nz": ref["grid"].get("nz"),
                      "Lx": ref["domain"]["Lx"][1], "Ly": ref["domain"]["Ly"][1],
                      "Lz": ref["domain"]["Lz"][1] if three else None,
                      "nt": ref["nt"], "ns": ref["num_snapshots"], "T": ref["domain"]["T"][1],
                      "m_type": ms[k % len(ms)], "c_type": cs[k % len(cs)]})
    out = descriptors(metas)
    assert list(out) == [name]
    got = out[name]
    assert list(got) == list(ref)
    for k in ref:
        if k != "num_trajectories":
            assert got[k] == ref[k], k


def _h5dump_header(path):
    return subprocess.run([H5DUMP, "-H", "-p", "-A", path], capture_output=True, text=True,
                          check=True).stdout


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
@pytest.mark.parametrize("cplx", [False, True])
def test_netcdf_class_file_layout_and_contents(h5, tmp_path, cplx):
    """Per-class NetCDF-4 file: the dimension scales and variables of the reference's
    aggregation (h5dump), and every trajectory's arrays at their offsets; a run with NaN
    is skipped and its slots keep netCDF's float fill value."""
    from nls.packaging import aggregate
    runs = tmp_path / "runs"
    runs.mkdir()
    ns, n, out_n = 6, 20, 8
    classes = [("constant", "layered"), ("defects", "constant"), ("phase", "waveguide")]
    paths = _write_runs(h5, str(runs), n, 5.0, 1.0, 60, ns, classes, out_n=out_n, cplx=cplx,
                        problem_type="cubic" if cplx else "klein_gordon", nan_at=1)
    res = aggregate(str(runs), str(tmp_path / "nc"), target_shape=(out_n, out_n))
    assert len(res) == 1
    (path, (ntraj, skipped)), = res.items()
    assert os.path.basename(path) == \
        f"class_2D_shape_{out_n}x{out_n}_ns_{ns}_{'complex' if cplx else 'real'}.nc"
    assert (ntraj, skipped) == (3, 1)
    hdr = _h5dump_header(path)
    spat, temp = out_n * out_n, ns * out_n * out_n
    for dim, length in (("total_temporal", 3 * temp), ("total_spatial", 3 * spat)):
        block = hdr[hdr.index(f'DATASET "{dim}"'):]
        block = block[:block.index("\n   }")]
        assert "DIMENSION_SCALE" in block and "_Netcdf4Dimid" in block
        assert f"This is a netCDF dimension but not a netCDF variable.{length:10d}" in block
        assert "H5T_IEEE_F32BE" in block
    tvars, svars = (("re_u", "im_u"), ("re_u0", "im_u0")) if cplx else (("u", "v"), ("u0", "v0"))
    for v in tvars + svars + ("m", "c"):
        block = hdr[hdr.index(f'DATASET "{v}"'):]
        block = block[:block.index("\n   }")]
        assert "H5T_IEEE_F32LE" in block and "DIMENSION_LIST" in block, v
        if v in tvars + svars:
            length = 3 * (temp if v in tvars else spat)
            assert f"CHUNKED ( {min(65536, length)} )" in block, v
        else:
            assert "CONTIGUOUS" in block, v
    for a in ("num_snapshots", "num_trajectories", "spatial_shape", "Lx", "Ly", "is_complex",
              "dims"):
        assert f'ATTRIBUTE "{a}"' in hdr, a
    assert h5.read(path, "u" if not cplx else "re_u").shape == (3 * temp,)
    # contents (m, c and the initial data resampled onto the stored grid, float32)
    from nls.launcher import downsample_interpolation
    fill = np.float32(9.9692099683868690e+36)
    rd = {v: h5.read(path, v) for v in tvars + svars + ("m", "c")}
    for t, p in enumerate(paths):
        ts, ss = slice(t * temp, (t + 1) * temp), slice(t * spat, (t + 1) * spat)
        if t == 1:  # the NaN run
            assert np.all(rd[tvars[0]][ts] == fill) and np.all(rd["m"][ss] == fill)
            continue
        u = h5.read(p, "u")
        want_t = (u.real, u.imag) if cplx else (u, h5.read(p, "v"))
        for v, w in zip(tvars, want_t):
            np.testing.assert_array_equal(rd[v][ts], w.astype(np.float32).ravel())

        def rs(a):
            a = np.asarray(a, np.float32)
            return downsample_interpolation(a[None], (out_n, out_n), 5.0, 5.0)[0]
        u0 = h5.read(p, "initial_condition/u0")
        want_s = (u0.real, u0.imag) if cplx else (u0, h5.read(p, "initial_condition/v0"))
        for v, w in zip(svars, want_s):
            np.testing.assert_allclose(rd[v][ss], rs(w).ravel(), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(rd["m"][ss], rs(h5.read(p, "focusing/m")).ravel(),
                                   rtol=1e-6, atol=1e-6)
        cname = "c" if cplx else "focusing/c"
        np.testing.assert_allclose(rd["c"][ss], rs(h5.read(p, cname)).ravel(), rtol=1e-6,
                                   atol=1e-6)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
@pytest.mark.parametrize("cplx", [False, True])
def test_canonical_class_file_layout_and_contents(h5, tmp_path, cplx, capsys):
    """canonicalize() on a class file and its run descriptor (canonicalize_data.py:39-262):
    coordinate variables member, time, x, y holding 0.., linspace(0, T, ns) and
    linspace(-L, L, n); u, v (u_real, u_imag) over (member, time, x, y) in one-trajectory
    chunks and m, c over (member, x, y), equal to the class file's flat slices reshaped;
    the descriptor's global attributes; the dry run prints the header and writes nothing."""
    from nls.packaging import aggregate, canonicalize, write_descriptors
    runs = tmp_path / "runs"
    runs.mkdir()
    ns, n, out_n, L, T = 5, 16, 6, 4.0, 2.5
    classes = [("constant", "layered"), ("defects", "constant"), ("phase", "waveguide")]
    _write_runs(h5, str(runs), n, L, T, 50, ns, classes, out_n=out_n, cplx=cplx,
                problem_type="cubic" if cplx else "klein_gordon")
    (cls, _), = aggregate(str(runs), str(tmp_path / "nc"), target_shape=(out_n, out_n)).items()
    cfg_path, = write_descriptors(str(runs), str(tmp_path / "cfg"))
    with open(cfg_path) as f:
        cfg = yaml.safe_load(f)
    assert canonicalize(cls, cfg_path, dry_run=True) is None
    hdr_txt = capsys.readouterr().out
    fields = ("u_real", "u_imag") if cplx else ("u", "v")
    assert f"float32 {fields[0]}(member, time, x, y) ;" in hdr_txt
    assert "float32 m(member, x, y) ;" in hdr_txt and "member = 3 ;" in hdr_txt
    assert not os.path.exists(cls.replace(".nc", "_standard.nc"))
    out = canonicalize(cls, cfg_path)
    assert out == cls.replace(".nc", "_standard.nc")
    hdr = _h5dump_header(out)
    for dim, length in (("member", 3), ("time", ns), ("x", out_n), ("y", out_n)):
        block = hdr[hdr.index(f'DATASET "{dim}"'):]
        block = block[:block.index("\n   }")]
        assert "DIMENSION_SCALE" in block and "_Netcdf4Dimid" in block, dim
        assert f"( {length} ) / ( {length} )" in block, dim
        assert ("H5T_STD_I32LE" if dim == "member" else "H5T_IEEE_F32LE") in block, dim
    for v in fields:
        block = hdr[hdr.index(f'DATASET "{v}"'):]
        block = block[:block.index("\n   }")]
        assert f"( 3, {ns}, {out_n}, {out_n} )" in block and "DIMENSION_LIST" in block, v
        assert f"CHUNKED ( 1, {ns}, {out_n}, {out_n} )" in block, v
    for v in ("m", "c"):
        block = hdr[hdr.index(f'DATASET "{v}"'):]
        block = block[:block.index("\n   }")]
        assert f"( 3, {out_n}, {out_n} )" in block and "CONTIGUOUS" in block, v
    # coordinates and data
    np.testing.assert_array_equal(h5.read(out, "member"), np.arange(3))
    np.testing.assert_allclose(h5.read(out, "time"), np.linspace(0, cfg["domain"]["T"][1], ns),
                               rtol=1e-6)
    np.testing.assert_allclose(h5.read(out, "x"), np.linspace(-L, L, out_n), rtol=1e-6)
    spat = out_n * out_n
    flat = ("re_u", "im_u") if cplx else ("u", "v")
    for fv, sv in zip(flat, fields):
        a, b = h5.read(cls, fv), h5.read(out, sv)
        for t in range(3):
            np.testing.assert_array_equal(b[t], a[t * ns * spat:(t + 1) * ns * spat]
                                          .reshape(ns, out_n, out_n))
    for v in ("m", "c"):
        a, b = h5.read(cls, v), h5.read(out, v)
        np.testing.assert_array_equal(b, a.reshape(3, out_n, out_n))
    # global attributes from the class file and the descriptor
    ra = lambda k: h5.read_attr(out, "/", k)
    assert ra("pde_type") == cfg["pde_type"] and ra("boundary_condition") == cfg["boundary_condition"]
    assert int(ra("num_trajectories")) == 3 and int(ra("num_snapshots")) == ns
    assert int(ra("nt")) == cfg["nt"] and ra("spatial_shape") == f"({out_n}, {out_n})"
    assert int(ra("dims")) == 2 and int(ra("is_complex")) == int(cplx)
    assert ra("Lx") == pytest.approx(L) and ra("T") == pytest.approx(cfg["domain"]["T"][1])
    assert ra("m_classes") == ", ".join(cfg["m_classes"])
    assert ra("c_classes") == ", ".join(cfg["c_classes"])
