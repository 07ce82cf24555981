"""Initial-condition samplers: the library's host draws (nls_ic_draw, used by the GPU
sampler) against the Python samplers' Philox parameter streams, and the closed forms the
Python fields follow.  Host only: no GPU is used."""
import numpy as np
import pytest

from nls.batch import ic_params
from nls.samplers import (ParamStream, bubbles_3d, gaussian_sum_2d, kink_ring_2d,
                          nlse_packets_2d, philox4x32_10)


def test_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10
    assert philox4x32_10([0, 0, 0, 0], (0, 0)).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                            0x9B00DBD8]
    assert philox4x32_10([0xFFFFFFFF] * 4, (0xFFFFFFFF, 0xFFFFFFFF)).tolist() == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


@pytest.mark.parametrize("seed", [0, 2, 123456789012345])
def test_draws_match_python_samplers(seed):
    for traj in (0, 1, 7, 99):
        _, _, g = gaussian_sum_2d(4, 10.0, seed, traj)
        assert np.array_equal(ic_params("gaussians_2d", seed, traj), np.array(g).ravel())
        _, _, k = kink_ring_2d(4, 10.0, seed, traj)
        assert np.array_equal(ic_params("kink_ring_2d", seed, traj), [k["R"], k["w"]])
        _, pk = nlse_packets_2d(4, 10.0, seed, traj)
        p = ic_params("packets_2d", seed, traj)
        assert p[0] == len(pk) and 2 <= p[0] <= 4
        assert np.allclose(p[1:].reshape(-1, 6), np.array(pk), rtol=1e-15, atol=0)
        # bubbles: the same stream, drawn here in the documented order
        ps = ParamStream(seed, traj, tag=0x3DB0)
        K = ps.integer(2, 5)
        ref = [K]
        for _ in range(K):
            ref += [ps.uniform(0.5, 2.0), ps.uniform(0.8, 2.0)] + list(ps.uniform(-4, 4, size=3))
        assert np.array_equal(ic_params("bubbles_3d", seed, traj), ref)


def test_draw_laws():
    """Parameter ranges of the configurations (ring radius U(2,6), width U(0.5,1.5); packet
    count in {2,3,4}, amplitude U(0.5,1.5), phase U(0, 2 pi); bubble count in {2..5})."""
    R = np.array([ic_params("kink_ring_2d", 2, t) for t in range(400)])
    assert R[:, 0].min() >= 2 and R[:, 0].max() <= 6 and R[:, 1].min() >= 0.5 and R[:, 1].max() <= 1.5
    assert abs(R[:, 0].mean() - 4.0) < 0.15
    K = [int(ic_params("packets_2d", 1, t)[0]) for t in range(300)]
    assert set(K) == {2, 3, 4}
    K3 = [int(ic_params("bubbles_3d", 3, t)[0]) for t in range(300)]
    assert set(K3) == {2, 3, 4, 5}


def test_python_fields_closed_forms():
    """The Python fields (the GPU sampler's checker) against their defining formulas at a
    few cells, evaluated independently of numpy's meshgrid."""
    n, L = 9, 10.0
    x = [-L + i * (2 * L / (n - 1)) for i in range(n)]
    u, v, prm = gaussian_sum_2d(n, L, 5, 2)
    for i, j in ((0, 0), (4, 7), (8, 3)):
        want = sum(A * np.exp(-((x[i] - cx) ** 2 + (x[j] - cy) ** 2) / w ** 2) for A, w, cx, cy in prm)
        assert abs(u[i, j] - want) < 1e-14
    assert not v.any()
    u, _, k = kink_ring_2d(n, L, 5, 2)
    assert abs(u[4, 4] - np.tanh(k["R"] / k["w"])) < 1e-15  # centre r = 0
    psi, pk = nlse_packets_2d(n, L, 5, 2)
    i, j = 3, 6
    want = sum(A * np.exp(-((x[i] - cx) ** 2 + (x[j] - cy) ** 2)) *
               np.exp(1j * (vx * (x[i] - cx) + vy * (x[j] - cy) + ph)) for A, ph, vx, vy, cx, cy in pk)
    assert abs(psi[i, j] - want) < 1e-14
    u3, _ = bubbles_3d(5, 8.0, 3, 1)
    assert u3.shape == (5, 5, 5) and u3.min() >= 0


def test_field_forms_match_reference_golden():
    """The kink ring and the coefficient-field classes whose form is a function of
    explicit parameters, evaluated at the parameters this library draws, against the
    reference's own functions at the same parameters (tests/golden/sampler_forms.npz,
    written by tests/golden/make_sampler_golden.py from real_sampler.py:435-467,
    c_fields_2d.py:12-52, m_fields_2d.py:14-82)."""
    import os
    from nls.fields import c_field, m_field
    from nls.samplers import kink_ring_2d
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "sampler_forms.npz"))
    n, L = int(g["n"]), float(g["L"])
    for s, t in g["seeds"]:
        s, t = int(s), int(t)
        tag = f"{s}_{t}"
        u, v, _ = kink_ring_2d(n, L, s, t)
        np.testing.assert_allclose(u, g[f"ring_u_{tag}"], rtol=0, atol=1e-15)
        np.testing.assert_allclose(v, g[f"ring_v_{tag}"], rtol=0, atol=0)
        for kind in ("constant", "periodic_structure", "piecewise_constant",
                     "sign_changing_mass"):
            np.testing.assert_allclose(c_field(kind, n, L, s, t)[0], g[f"c_{kind}_{tag}"],
                                       rtol=1e-15, atol=1e-15, err_msg=kind)
        c = g[f"c_for_m_{tag}"]
        np.testing.assert_array_equal(c, c_field("sign_changing_mass", n, L, s, t)[0])
        for kind in ("constant", "piecewise", "gradient", "phase", "topological"):
            np.testing.assert_allclose(m_field(kind, n, L, s, t, c=c)[0], g[f"m_{kind}_{tag}"],
                                       rtol=1e-14, atol=1e-14, err_msg=kind)


def _golden():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "sampler_forms.npz"))


def test_random_field_classes_match_reference_by_replay():
    """The seven field classes whose draws happen inside the reference functions (m:
    defects, quasiperiodic, multiscale; c: layered, waveguide, quasiperiodic, turbulent)
    against the reference's generate_m_fields / generate_c_fields run with this library's
    draws replayed into np.random (tests/golden/make_sampler_golden.py): the same forms
    from the same draws in the same order."""
    from nls.fields import c_field, m_field
    g = _golden()
    n, L = int(g["n"]), float(g["L"])
    for s, t in g["seeds"]:
        s, t = int(s), int(t)
        tag = f"{s}_{t}"
        for kind in ("defects", "quasiperiodic", "multiscale"):
            np.testing.assert_allclose(m_field(kind, n, L, s, t)[0], g[f"mr_{kind}_{tag}"],
                                       rtol=1e-14, atol=1e-14, err_msg=kind)
        for kind in ("layered", "waveguide", "quasiperiodic", "turbulent"):
            np.testing.assert_allclose(c_field(kind, n, L, s, t)[0], g[f"cr_{kind}_{tag}"],
                                       rtol=1e-14, atol=1e-14, err_msg=kind)


def test_reference_phenomena_match_by_replay():
    """kink_field (sine-Gordon kinks and the anisotropic GRF velocity) and
    multi_soliton_state (every system type, arrangement and phase pattern) against the
    reference's own methods run with this library's draws replayed, and
    fundamental_soliton at explicit parameters for the four system types."""
    from nls import samplers as S
    g = _golden()
    n, L = int(g["replay_n"]), float(g["replay_L"])
    for s, t in g["replay_seeds"]:
        s, t = int(s), int(t)
        tag = f"{s}_{t}"
        u, v, _ = S.kink_field_2d(n, L, s, t)
        np.testing.assert_allclose(u, g[f"kink_u_{tag}"], rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(v, g[f"kink_v_{tag}"], rtol=1e-12, atol=1e-13)
        psi, _ = S.multi_soliton_2d(n, L, s, t, normalize=False)
        np.testing.assert_allclose(psi, g[f"solitons_{tag}"], rtol=1e-13, atol=1e-14)
    cases = [("cubic", "circular", "alternating"), ("cubic_quintic", "linear", "synchronized"),
             ("saturable", "lattice", "random"), ("glasner_allen_flowers", "random", "vortex"),
             ("cubic", "random", "partial_coherence")]
    for k, (st, arr, ph) in enumerate(cases):
        psi, _ = S.multi_soliton_2d(n, L, 9, k, system_type=st, arrangement=arr,
                                    phase_pattern=ph, velocity_scale=0.5, normalize=False)
        np.testing.assert_allclose(psi, g[f"solitons_case{k}"], rtol=1e-13, atol=1e-14)
    X, Y = S.grid_2d(n, L)
    for k, st in enumerate(("cubic", "cubic_quintic", "saturable", "glasner_allen_flowers")):
        kw = dict(amplitude=1.3, width=0.9, position=(0.7, -1.1), phase=0.4,
                  velocity=(0.3, -0.6), Lambda=0.08, chirp_factor=0.05, aspect_ratio=1.3,
                  orientation=0.6, order=1 + (k % 2))
        np.testing.assert_allclose(S.fundamental_soliton(X, Y, L, st, **kw), g[f"soliton_{st}"],
                                   rtol=1e-14, atol=1e-15)


# the reference's dataset domains (config_*.yaml: KGE 2D and 3D on [-3, 3], NLSE 2D on
# [-10, 10]) and the configurations' own (side 20 in 2D, 16 in 3D)
@pytest.mark.parametrize("L", [3.0, 10.0])
def test_samplers_place_structures_inside_the_domain(L):
    """On every domain the sampled structure lies inside it: Gaussian sums peak above 0.1
    with centres inside, kink rings cross zero inside the grid (so the ring is not wholly
    outside), packets and bubbles have their centres inside, and the reference-form
    phenomena are finite and non-trivial."""
    from nls import samplers as S
    n = 48
    for t in range(60):
        u, _, g = S.gaussian_sum_2d(n, L, 2, t)
        assert u.max() > 0.1
        assert all(abs(cx) < L and abs(cy) < L for _, _, cx, cy in g)
        u, _, k = S.kink_ring_2d(n, L, 2, t)
        assert u.min() < 0 < u.max() and k["R"] < L, (t, k)
        _, pk = S.nlse_packets_2d(n, L, 1, t)
        assert all(abs(p[4]) < L and abs(p[5]) < L for p in pk)
    for t in range(10):
        u, _ = S.bubbles_3d(16, L, 3, t)
        assert u.max() > 0.1
        ub, vb, _ = S.kink_field_2d(n, L, 4, t)
        assert np.all(np.isfinite(ub)) and np.all(np.isfinite(vb))
        psi, _ = S.multi_soliton_2d(n, L, 4, t)
        assert np.all(np.isfinite(psi)) and abs(np.abs(psi).max() - 1.0) < 1e-12


@pytest.mark.parametrize("L", [3.0, 8.0, 10.0])
def test_draws_match_python_samplers_any_domain(L):
    for traj in (0, 5):
        _, _, g = gaussian_sum_2d(4, L, 7, traj)
        assert np.array_equal(ic_params("gaussians_2d", 7, traj, L=L), np.array(g).ravel())
        _, _, k = kink_ring_2d(4, L, 7, traj)
        assert np.array_equal(ic_params("kink_ring_2d", 7, traj, L=L), [k["R"], k["w"]])
        _, pk = nlse_packets_2d(4, L, 7, traj)
        p = ic_params("packets_2d", 7, traj, L=L)
        assert np.allclose(p[1:].reshape(-1, 6), np.array(pk), rtol=1e-15, atol=0)
        ps = ParamStream(7, traj, tag=0x3DB0)
        K = ps.integer(2, 5)
        ref = [K]
        b = 4.0 * L / 8.0
        for _ in range(K):
            ref += [ps.uniform(0.5, 2.0), ps.uniform(0.8, 2.0)] + list(ps.uniform(-b, b, size=3))
        assert np.array_equal(ic_params("bubbles_3d", 7, traj, L=L), ref)
