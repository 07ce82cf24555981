"""End-to-end dataset generation: dataset YAML -> fused batch on the GPU -> HDF5 files
in the launchers' layout, checked against the FP64 oracle on the same inputs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def h5():
    from nls import h5 as mod
    if mod.available() is None:
        pytest.skip("libhdf5 not loadable")
    return mod


@pytest.mark.parametrize("integrator", ["gautschi", "sv"])
def test_launcher_kge_dataset_matches_oracle(orc, h5, tmp_path, integrator):
    """Gautschi (FP32 Krylov, the launcher's full-T Lanczos against the oracle's) within the
    1e-5 bar; Stormer-Verlet (FP64, the reference's arithmetic) within 1e-6: the batch keeps
    m and c in FP32, the oracle reads the files' FP64 fields."""
    from nls.launcher import BatchLauncher, load_config
    cfg = load_config(os.path.join(GOLDEN, "dataset_kge_2d.yaml"), output_dir=str(tmp_path),
                      run_id="t", dr_strategy="none", batch=2, integrator=integrator)
    res = BatchLauncher(cfg).run(log=lambda *a: None)
    assert res["files"] == 3
    for idx in range(3):
        p = str(tmp_path / "hdf5" / f"run_t_{idx:04d}.h5")
        u, v = h5.read(p, "u"), h5.read(p, "v")
        u0, v0 = h5.read(p, "initial_condition/u0"), h5.read(p, "initial_condition/v0")
        m, c = h5.read(p, "focusing/m"), h5.read(p, "focusing/c")
        assert u.shape == (5, 48, 48) and v.shape == (5, 48, 48)
        ur, _ = orc.kge_run(48, 10.0, 0.5, 100, 5, m, c, u0, v0, krylov=10, mu2=1.0,
                            method=1 if integrator == "sv" else 0, variant=1)
        for k in range(5):
            assert rel_l2(u[k], ur[k]) < (1e-6 if integrator == "sv" else 1e-5), (idx, k)
        assert h5.read_attr(p, "metadata", "run_index") == idx
        assert h5.read_attr(p, "focusing", "type") == cfg.m_classes[(idx // 2) % 2]


def test_launcher_nlse_phi4_and_downsampling(h5, tmp_path):
    from nls.launcher import BatchLauncher, LaunchConfig
    for pde, name, ph in (("Nonlinear Schrodinger", "n", "multi_soliton"),
                          ("phi4", "p", "kink_field")):
        cfg = LaunchConfig(pde_type=pde, num_runs=2, num_snapshots=4, nt=40, T=0.2, nx=32, ny=32,
                           Lx=8.0, Ly=8.0, output_dir=str(tmp_path / name), run_id=name,
                           krylov_dim=12, dr_strategy="interpolation", dr_x=16, dr_y=16,
                           phenomenon=ph)
        res = BatchLauncher(cfg).run(log=lambda *a: None)
        assert res["files"] == 2
        p = str(tmp_path / name / "hdf5" / f"run_{name}_0001.h5")
        assert h5.read_attr(p, "metadata", "phenomenon") == ph
        u = h5.read(p, "u")
        assert u.shape == (4, 16, 16) and np.all(np.isfinite(u))
        if name == "n":
            assert np.iscomplexobj(u) and h5.read(p, "c").shape == (32, 32)
        else:
            assert h5.read(p, "v").shape == (4, 16, 16)
