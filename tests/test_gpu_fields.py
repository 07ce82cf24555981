"""Coefficient fields and reference phenomena sampled on the GPU (nls_field_sample,
nls_phenomenon_sample) against the reference's own functions at the same draws
(tests/golden/sampler_forms.npz, written by tests/golden/make_sampler_golden.py from
c_fields_2d.py, m_fields_2d.py, real_sampler.py and nlse_sampler.py) and against the
host generators of the package on larger grids.  Both sides compute in FP64; they differ
by the last bits of the transcendental functions and by the summation order of the
Gaussian filters, the DFTs and the reductions, hence the 1e-12 bars."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C_CLASSES = ("constant", "periodic_structure", "piecewise_constant", "sign_changing_mass",
             "layered", "waveguide", "quasiperiodic", "turbulent")
M_CLASSES = ("constant", "piecewise", "gradient", "phase", "topological", "defects",
             "quasiperiodic", "multiscale")


def _golden():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "sampler_forms.npz"))


def close(a, b, tol=1e-12):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) <= tol * scale


def test_fields_on_gpu_match_reference_golden():
    """All 16 classes at the fixtures' draws: the closed forms (c: constant, periodic,
    piecewise, sign-changing; m: constant, piecewise, gradient, phase, topological, the
    last three from the fixture's c field) and the seven replayed classes."""
    from nls.device import DeviceArray
    from nls.fields import c_field_device, m_field_device
    g = _golden()
    n, L = int(g["n"]), float(g["L"])
    for s, t in g["seeds"]:
        s, t = int(s), int(t)
        tag = f"{s}_{t}"
        print(f"seed {s} traj {t}", flush=True)
        for kind in ("constant", "periodic_structure", "piecewise_constant",
                     "sign_changing_mass"):
            assert close(c_field_device(kind, n, L, s, t).to_host(), g[f"c_{kind}_{tag}"]), kind
        c = DeviceArray.from_host(g[f"c_for_m_{tag}"])
        for kind in ("constant", "piecewise", "gradient", "phase", "topological"):
            got = m_field_device(kind, n, L, s, t, c=c).to_host()
            assert close(got, g[f"m_{kind}_{tag}"]), kind
        for kind in ("defects", "quasiperiodic", "multiscale"):
            assert close(m_field_device(kind, n, L, s, t).to_host(), g[f"mr_{kind}_{tag}"]), kind
        for kind in ("layered", "waveguide", "quasiperiodic", "turbulent"):
            assert close(c_field_device(kind, n, L, s, t).to_host(), g[f"cr_{kind}_{tag}"]), kind


@pytest.mark.parametrize("n,L", [(128, 10.0), (97, 3.0)])
def test_fields_on_gpu_match_host_generators(n, L):
    """Every class on grids beyond the fixtures' (odd n included) against nls/fields.py,
    several trajectories each; m classes derived from c take a non-trivial c field."""
    from nls.fields import c_field, c_field_device, m_field, m_field_device
    for t in range(3):
        print(f"n {n} traj {t}", flush=True)
        for kind in C_CLASSES:
            ref = c_field(kind, n, L, 7, t)[0]
            assert close(c_field_device(kind, n, L, 7, t).to_host(), ref), (kind, t)
        cd = c_field_device("periodic_structure", n, L, 7, t)
        ch = c_field("periodic_structure", n, L, 7, t)[0]
        for kind in M_CLASSES:
            ref = m_field(kind, n, L, 7, t, c=ch)[0]
            assert close(m_field_device(kind, n, L, 7, t, c=cd).to_host(), ref), (kind, t)
        # no c field: the derived classes are constant, as in the reference's dispatch
        assert np.all(m_field_device("gradient", n, L, 7, t).to_host() == 1.0)


def test_phenomena_on_gpu_match_reference_golden():
    """kink_field (kinks and the anisotropic random velocity field) and multi_soliton_state
    at the replay fixtures' seeds."""
    from nls.samplers import kink_field_device, multi_soliton_device
    g = _golden()
    n, L = int(g["replay_n"]), float(g["replay_L"])
    for s, t in g["replay_seeds"]:
        s, t = int(s), int(t)
        tag = f"{s}_{t}"
        print(f"seed {s} traj {t}", flush=True)
        u, v, _ = kink_field_device(n, L, s, t)
        assert close(u.to_host(), g[f"kink_u_{tag}"])
        assert close(v.to_host(), g[f"kink_v_{tag}"], 1e-11)
        psi, _ = multi_soliton_device(n, L, s, t, normalize=False)
        assert close(psi.to_host(), g[f"solitons_{tag}"])


def test_phenomena_on_gpu_match_host_samplers():
    """40 trajectories of each phenomenon at n = 96 against the host samplers: every
    arrangement, phase pattern and system type the space draws, and the same parameters."""
    from nls import samplers as S
    n, L = 96, 10.0
    seen = set()
    for t in range(40):
        u, v, prm = S.kink_field_device(n, L, 5, t)
        ur, vr, prm_r = S.kink_field_2d(n, L, 5, t)
        assert prm == prm_r, t
        assert close(u.to_host(), ur), t
        assert close(v.to_host(), vr, 1e-11), t
        psi, p = S.multi_soliton_device(n, L, 5, t)
        psi_r, p_r = S.multi_soliton_2d(n, L, 5, t)
        for k in ("system_type", "phase_pattern", "arrangement", "n_solitons"):
            assert p[k] == p_r[k], (t, k)
        assert close(psi.to_host(), psi_r), t
        assert abs(np.max(np.abs(psi.to_host())) - 1.0) < 1e-15
        seen.add((p["arrangement"], p["phase_pattern"]))
    assert len({a for a, _ in seen}) == 4 and len({b for _, b in seen}) == 4


def test_field_sampler_rejects_bad_arguments():
    from nls._lib import NLSError
    from nls.fields import c_field_device, m_field_device
    from nls.samplers import _phenomenon_device
    with pytest.raises(NLSError, match="unknown c class"):
        c_field_device("marble", 16, 1.0, 0, 0)
    with pytest.raises(NLSError, match="unknown m class"):
        m_field_device("marble", 16, 1.0, 0, 0)
    with pytest.raises(NLSError, match="n must lie"):
        c_field_device("constant", 1, 1.0, 0, 0)
    with pytest.raises(NLSError, match="unknown phenomenon"):
        _phenomenon_device("breather", 16, 1.0, 0, 0, False, False)
