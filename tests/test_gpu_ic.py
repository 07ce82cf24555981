"""GPU initial-condition samplers (nls_batch_set_initial_sampled) against the Python
samplers on the same (seed, trajectory): the fields agree to the last bits of exp, tanh,
sin and cos; the trajectories then evolve as with host-supplied initial data."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(a, b, tol):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) < tol


@pytest.mark.parametrize("kind", ["gaussians_2d", "kink_ring_2d"])
def test_sampled_kge_2d_matches_python(kind):
    from nls.batch import FusedBatch
    from nls.samplers import gaussian_sum_2d, kink_ring_2d
    fn = {"gaussians_2d": gaussian_sum_2d, "kink_ring_2d": kink_ring_2d}[kind]
    ns, L, seed = [33, 64, 130], 10.0, 11
    kw = dict(krylov_dim=8, mu2=1.0, cells_per_thread=16, pass_kernel=2, lanes=1,
              snapshot_every=5, max_snapshots=3)
    a = FusedBatch(ns, L, 0.005, **kw)
    b = FusedBatch(ns, L, 0.005, **kw)
    for t, n in enumerate(ns):
        a.sample_initial(t, kind, seed, traj_index=10 + t, m=1.0, c=1.0)
        u0, v0, _ = fn(n, L, seed, 10 + t)
        b.set_initial(t, np.ones((n, n)), np.ones((n, n)), u0, v0)
        ug, vg = a.snapshots(t)
        assert ug.shape[0] == 1 and close(ug[0], u0, 1e-14) and not vg[0].any(), (kind, n)
        assert close(a.state(t)[0], u0, 1e-14)
    a.run(10)
    b.run(10)
    for t in range(len(ns)):
        assert close(a.state(t)[0], b.state(t)[0], 1e-12)
        assert close(a.snapshots(t)[0][2], b.snapshots(t)[0][2], 1e-12)


def test_sampled_nlse_packets_match_python():
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    ns, L, seed = [40, 64], 10.0, 1
    b = FusedBatch(ns, L, 0.005, equation="nlse", method="strang", krylov_dim=12,
                   cells_per_thread=8, pass_kernel=0, snapshot_every=4, max_snapshots=2)
    h = FusedBatch(ns, L, 0.005, equation="nlse", method="strang", krylov_dim=12,
                   cells_per_thread=8, pass_kernel=0, snapshot_every=4, max_snapshots=2)
    for t, n in enumerate(ns):
        b.sample_initial(t, "packets_2d", seed)
        psi, _ = nlse_packets_2d(n, L, seed, t)
        h.set_initial(t, np.ones((n, n)), np.ones((n, n)), psi)
        s0, _ = b.snapshots(t)
        assert close(s0[0], psi, 1e-14)
        assert close(b.state(t)[0], psi, 1e-7)  # complex64 state
    b.run(4)
    h.run(4)
    for t in range(len(ns)):
        assert close(b.snapshots(t)[0][1], h.snapshots(t)[0][1], 1e-6)


def test_sampled_bubbles_3d_match_python():
    from nls.batch import FusedBatch
    from nls.samplers import bubbles_3d
    ns, L, seed = [20, 33], 8.0, 3
    b = FusedBatch(ns, L, 0.004, method="sewi", dim=3, krylov_dim=10, mu2=1.0,
                   cells_per_thread=4, pass_kernel=4)
    for t, n in enumerate(ns):
        b.sample_initial(t, "bubbles_3d", seed, m=1.0, c=1.0)
        u0, _ = bubbles_3d(n, L, seed, t)
        assert close(b.state(t)[0], u0, 1e-14)


def test_sampler_kind_checks():
    from nls._lib import NLSError
    from nls.batch import FusedBatch
    b = FusedBatch([32], 10.0, 0.005, krylov_dim=8)
    for kind in ("packets_2d", "bubbles_3d"):
        with pytest.raises(NLSError):
            b.sample_initial(0, kind, 1)
    b.sample_initial(0, "kink_ring_2d", 1)
    b.run(1)
    with pytest.raises(NLSError):  # initial data only before the first step
        b.sample_initial(0, "kink_ring_2d", 1)
