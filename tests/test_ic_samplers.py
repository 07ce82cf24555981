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
