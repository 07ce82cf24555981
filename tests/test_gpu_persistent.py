"""GPU tests of the persistent step kernel (nls_batch_params.persistent): one cooperative
launch per run of steps without snapshots, per-trajectory arrival counters between the
phases of a step.  Its results must be bitwise those of the graph-captured steps at the
same tile (same kernels' arithmetic, same reduction order), a trajectory's bits must not
depend on its batch-mates, and it must match the FP64 oracle."""
import numpy as np
import pytest

from test_gpu_fused import TOL, make_case, rel_l2

pytestmark = pytest.mark.gpu


def _run(ns_list, cases, steps, persistent, variant, reorth=0, uniform=False, cpt=16,
         krylov=8, every=0, nsnap=0, chunks=(1.0,)):
    from nls.batch import FusedBatch
    L, dt = 10.0, 0.005
    b = FusedBatch(ns_list, L, dt, krylov_dim=krylov, mu2=1.0, cells_per_thread=cpt,
                   pass_kernel=2, variant=variant, reorth=reorth, persistent=persistent,
                   snapshot_every=every, max_snapshots=nsnap)
    for t, (m, c, u0, v0) in enumerate(cases):
        if uniform:
            m, c = np.ones_like(m), np.ones_like(c)
        b.set_initial(t, m, c, u0, v0)
    done = 0
    for f in chunks:  # several runs: the state ring position carries over between launches
        k = int(round(f * steps)) - done
        b.run(k)
        done += k
    snaps = [b.snapshots(t)[0] if nsnap else None for t in range(len(ns_list))]
    return b, [b.state(t)[0] for t in range(len(ns_list))], snaps


@pytest.mark.parametrize("variant", ["full", "reference"])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("reorth", [0, 2])
def test_persistent_equals_graph_steps_bitwise(variant, uniform, reorth):
    ns_list = [64, 40, 96, 130, 33]
    cases = [make_case(n, 8, t, ring=(t == 1)) for t, n in enumerate(ns_list)]
    steps, every, nsnap = 61, 20, 4
    kw = dict(variant=variant, reorth=reorth, uniform=uniform, every=every, nsnap=nsnap,
              chunks=(0.4, 1.0))
    bp, up, sp = _run(ns_list, cases, steps, True, **kw)
    _, ug, sg = _run(ns_list, cases, steps, False, **kw)
    # steps 20, 40, 60 take snapshots on the captured steps; the other 58 run persistent
    assert bp.persistent_steps() == steps - 3
    for t in range(len(ns_list)):
        assert np.array_equal(up[t], ug[t]), t
        assert np.array_equal(sp[t], sg[t]), t
        assert np.all(np.isfinite(up[t]))


def test_persistent_alone_equals_in_mix_bitwise():
    mix = [64, 40, 96, 64, 300]
    target = 2
    cases = [make_case(n, 4, t) for t, n in enumerate(mix)]
    _, u_mix, _ = _run(mix, cases, 40, True, "full")
    _, u_alone, _ = _run([mix[target]], [cases[target]], 40, True, "full")
    assert np.array_equal(u_mix[target], u_alone[0])


@pytest.mark.parametrize("variant,krylov", [("full", 8), ("reference", 10)])
def test_persistent_vs_oracle(orc, variant, krylov):
    from nls.batch import FusedBatch
    ns_list = [128, 64]
    L, dt, nt, ns = 10.0, 0.005, 301, 3
    cases = [make_case(n, 2, t, ring=(t == 0)) for t, n in enumerate(ns_list)]
    b = FusedBatch(ns_list, L, dt, krylov_dim=krylov, mu2=1.0, snapshot_every=nt // ns,
                   max_snapshots=ns, cells_per_thread=16, pass_kernel=2, variant=variant,
                   persistent=True)
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(nt - 1)
    assert b.persistent_steps() == nt - 1 - (ns - 1)
    code = 1 if variant == "full" else 0
    for t, n in enumerate(ns_list):
        m, c, u0, v0 = cases[t]
        ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=krylov, mu2=1.0,
                            variant=code)
        u, _ = b.snapshots(t)
        for k in range(ns):
            assert rel_l2(u[k], ur[k]) < TOL, (t, k)


def test_persistent_automatic_tile_fits_and_matches_graph_steps():
    """Without a tile the persistent kernel takes the smallest whose chunks fit the
    resident workgroups; the captured steps at that tile give the same bits."""
    ns_list = [300] * 12 + [256, 192]
    cases = [make_case(n, 3, t, ring=True) for t, n in enumerate(ns_list)]
    bp, up, _ = _run(ns_list, cases, 30, True, "full", uniform=True, cpt=0)
    tile = bp.tile()
    assert tile["pass_kind"] == 2 and 4 <= tile["cells_per_thread"] <= 64
    assert bp.persistent_steps() == 30
    _, ug, _ = _run(ns_list, cases, 30, False, "full", uniform=True,
                    cpt=tile["cells_per_thread"])
    for t in range(len(ns_list)):
        assert np.array_equal(up[t], ug[t]), t


def test_persistent_reorth_switch_hands_over_to_graph_steps():
    """A 6 x 6 grid with m = 24 exhausts its Krylov space and loses orthogonality: in
    mode 2 the host sees the device flag at a 16-step check and runs the
    reorthogonalising captured steps from then on, as without the persistent kernel."""
    ns_list = [6, 40]
    cases = [make_case(n, 9, t) for t, n in enumerate(ns_list)]
    bp, up, _ = _run(ns_list, cases, 48, True, "full", reorth=2, krylov=24, cpt=8)
    _, ug, _ = _run(ns_list, cases, 48, False, "full", reorth=2, krylov=24, cpt=8)
    assert 0 < bp.persistent_steps() < 48
    assert bp.monitor(0)["reorth_on"]
    for t in range(len(ns_list)):
        assert np.array_equal(up[t], ug[t]), t


@pytest.mark.parametrize("kw", [dict(dim=3), dict(equation="nlse", method="strang"),
                                dict(reorth=1), dict(pass_kernel=1, cells_per_thread=4)])
def test_persistent_refuses_unsupported_batches(kw):
    from nls._lib import NLSError
    from nls.batch import FusedBatch
    n = 24
    args = dict(krylov_dim=8, persistent=True)
    args.update(kw)
    b = FusedBatch([n], 10.0, 0.005, **args)
    shape = (n,) * b.dim
    if b.equation == "nlse":
        b.set_initial(0, np.ones(shape), np.ones(shape), np.zeros(shape, np.complex128))
    else:
        b.set_initial(0, np.ones(shape), np.ones(shape), np.zeros(shape), np.zeros(shape))
    with pytest.raises(NLSError, match="persistent"):
        b.run(2)


def test_persistent_abort_flag_ends_the_launch():
    """The host sets the abort flag before a run: every wait of the persistent kernel that
    finds its counter short reads the flag and gives up, every workgroup exits, and the run
    fails with a message.  A new persistent batch afterwards runs normally (bitwise the
    captured steps)."""
    from nls._lib import NLSError
    ns_list = [64, 96]
    cases = [make_case(n, 8, t) for t, n in enumerate(ns_list)]
    from nls.batch import FusedBatch
    b = FusedBatch(ns_list, 10.0, 0.005, krylov_dim=8, mu2=1.0, cells_per_thread=16,
                   pass_kernel=2, variant="full", persistent=True)
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.abort()
    with pytest.raises(NLSError, match="aborted by the host"):
        b.run(8)
    b.close()
    _, up, _ = _run(ns_list, cases, 20, True, "full")
    _, ug, _ = _run(ns_list, cases, 20, False, "full")
    for t in range(len(ns_list)):
        assert np.array_equal(up[t], ug[t]), t
