"""Coefficient-field classes, downsampling and dataset-config parsing of the
launcher (host-side parts; the GPU run is in test_gpu_launcher.py)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("kind", ["constant", "periodic_structure", "piecewise_constant",
                                  "sign_changing_mass", "layered", "waveguide", "quasiperiodic",
                                  "turbulent"])
def test_c_classes_finite_and_reproducible(kind):
    from nls.fields import c_field
    a, pa = c_field(kind, 40, 5.0, 3, 11)
    b, pb = c_field(kind, 40, 5.0, 3, 11)
    assert a.shape == (40, 40) and np.all(np.isfinite(a))
    assert np.array_equal(a, b) and pa == pb
    if kind not in ("constant", "sign_changing_mass"):
        assert not np.array_equal(a, c_field(kind, 40, 5.0, 3, 12)[0])


@pytest.mark.parametrize("kind", ["constant", "piecewise", "gradient", "phase", "topological",
                                  "defects", "quasiperiodic", "multiscale"])
def test_m_classes_finite_and_reproducible(kind):
    from nls.fields import c_field, m_field
    c, _ = c_field("layered", 40, 5.0, 1, 2)
    a, _ = m_field(kind, 40, 5.0, 1, 2, c=c)
    b, _ = m_field(kind, 40, 5.0, 1, 2, c=c)
    assert a.shape == (40, 40) and np.all(np.isfinite(a)) and np.array_equal(a, b)
    if kind in ("defects", "quasiperiodic", "multiscale"):
        assert a.min() >= 0.1 - 1e-15  # floor at 0.1 m0


def test_closed_form_classes():
    from nls.fields import c_field
    n, L = 33, 4.0
    x = np.linspace(-L, L, n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    p, prm = c_field("periodic_structure", n, L, 0, 0)
    k = np.pi * prm["frequency"] / L
    np.testing.assert_allclose(p, 1 + prm["amplitude"] * np.sin(k * X) * np.sin(k * Y), atol=1e-14)
    q, prm = c_field("piecewise_constant", n, L, 0, 0)
    assert set(np.unique(q)) <= {1.0, prm["contrast_factor"]}
    assert np.all(q[X < -L + 2 * L / prm["num_layers"]] == 1.0)  # first layer keeps the base


def test_interpolation_downsampling_matches_scipy():
    from scipy.interpolate import RegularGridInterpolator
    from nls.launcher import downsample_interpolation
    rng = np.random.default_rng(0)
    u = rng.standard_normal((3, 50, 50))
    L = 7.0
    out = downsample_interpolation(u, (17, 23), L, L)
    x = np.linspace(-L, L, 50)
    Xn, Yn = np.meshgrid(np.linspace(-L, L, 17), np.linspace(-L, L, 23), indexing="ij")
    pts = np.vstack([Xn.ravel(), Yn.ravel()]).T
    for t in range(3):
        ref = RegularGridInterpolator((x, x), u[t], method="linear", bounds_error=False,
                                      fill_value=None)(pts).reshape(17, 23)
        np.testing.assert_allclose(out[t], ref, rtol=0, atol=1e-13)


def test_fft_downsampling_keeps_band_limited_modes():
    """A trigonometric polynomial below the coarse Nyquist frequency is kept exactly
    up to the orthonormal scaling n_fine / n_coarse."""
    from nls.launcher import downsample_fft
    nf, nc = 64, 16
    j = np.arange(nf)
    f = (np.cos(2 * np.pi * 3 * j / nf)[:, None] * np.sin(2 * np.pi * 2 * j / nf)[None, :])[None]
    out = downsample_fft(f, (nc, nc))
    jc = np.arange(nc)
    g = np.cos(2 * np.pi * 3 * jc / nc)[:, None] * np.sin(2 * np.pi * 2 * jc / nc)[None, :]
    np.testing.assert_allclose(out[0], g * nf / nc, atol=1e-13)
    z = downsample_fft(f.astype(np.complex128), (nc, nc))
    assert np.iscomplexobj(z)


def test_yaml_config_keys():
    from nls.launcher import load_config
    c = load_config(os.path.join(GOLDEN, "dataset_kge_2d.yaml"), output_dir="x")
    assert c.pde_type == "Klein-Gordon" and c.num_runs == 3 and c.num_snapshots == 5
    assert c.nt == 100 and c.T == 0.5 and c.Lx == 10.0 and c.nx == c.ny == 48
    assert c.m_classes == ["constant", "defects"] and c.c_classes == ["constant", "layered"]
    assert c.krylov_dim == 10 and c.mu2 == 1.0 and c.seed == 7 and c.output_dir == "x"
    assert c.dr_strategy == "interpolation" and c.dr_x == 32


def test_launcher_rejects_unbuilt_equations(tmp_path):
    from nls.launcher import BatchLauncher, LaunchConfig
    with pytest.raises(ValueError, match="integrator"):
        BatchLauncher(LaunchConfig(pde_type="sine_gordon", output_dir=str(tmp_path)))


def test_launcher_refuses_unbuilt_phenomena_and_labels_what_it_generates(tmp_path):
    """A phenomenon that is not built is refused (nothing else is generated under its
    name); the default is the reference launchers' (kink_field for real fields,
    multi_soliton for NLSE), and 3D refuses coefficient-field classes it does not build."""
    from nls.launcher import BatchLauncher, LaunchConfig, resolve_phenomenon
    assert resolve_phenomenon("", "real2d") == "kink_field"
    assert resolve_phenomenon("", "nlse2d") == "multi_soliton"
    assert resolve_phenomenon("gaussian", "real2d") == "gaussians"
    for bad, fam in (("ring_soliton", "real2d"), ("kink_solution", "real2d"),
                     ("kink_field", "nlse2d"), ("vortex_lattice", "nlse2d"),
                     ("kink_field", "real3d")):
        with pytest.raises(ValueError, match="not built"):
            resolve_phenomenon(bad, fam)
    with pytest.raises(ValueError, match="not built"):
        BatchLauncher(LaunchConfig(pde_type="klein_gordon", phenomenon="breather_field",
                                   output_dir=str(tmp_path)))
    with pytest.raises(ValueError, match="3D coefficient"):
        BatchLauncher(LaunchConfig(pde_type="klein_gordon", dim=3, nx=16, ny=16, nz=16,
                                   m_classes=["defects"], output_dir=str(tmp_path)))
    lk = BatchLauncher(LaunchConfig(pde_type="klein_gordon", nx=24, ny=24, Lx=3.0, Ly=3.0,
                                    output_dir=str(tmp_path), sampler="host"))
    u0, v0, m, c, prm, _, _ = lk.sample(0)
    assert lk.phenomenon == "kink_field" and "winding_x" in prm and np.any(v0 != 0)
    ln = BatchLauncher(LaunchConfig(pde_type="Nonlinear Schrodinger", nx=24, ny=24,
                                    output_dir=str(tmp_path), sampler="host"))
    psi, _, _, _, prm, _, _ = ln.sample(1)
    assert ln.phenomenon == "multi_soliton" and "n_solitons" in prm
    assert np.iscomplexobj(psi) and abs(np.abs(psi).max() - 1.0) < 1e-12
