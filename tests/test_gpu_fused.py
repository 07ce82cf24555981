"""GPU parity of the fused batched FP32/FP64 Klein-Gordon path against the FP64
oracle (reference Lanczos variant), plus determinism properties: a trajectory's
result does not depend on its batch-mates, graphs replay the eager result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# relative-L2 bar of the fused FP32 Krylov / FP64 state path vs the FP64 oracle
TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make_case(n, seed, traj, L=10.0, ring=False):
    from nls.samplers import gaussian_sum_2d, kink_ring_2d
    rng = np.random.default_rng(seed * 100 + traj)
    if ring:
        u0, v0, _ = kink_ring_2d(n, L, seed, traj)
    else:
        u0, v0, _ = gaussian_sum_2d(n, L, seed, traj)
    m = 0.8 + 0.4 * rng.random((n, n))
    c = 0.8 + 0.4 * rng.random((n, n))
    return m, c, u0, v0


# (tile, pass kernel): LDS-tiled chunk, row march, row march storing only the basis
# (384^2 takes its window-shift path through registers, 3n > 1024)
@pytest.mark.parametrize("cpt,kernel", [(4, 1), (32, 2), (16, 3), (64, 3)])
def test_fused_kge_uneven_batch_vs_oracle(orc, cpt, kernel):
    from nls.batch import FusedBatch
    ns_list = [48, 64, 33, 130] + ([384] if kernel == 3 else [])
    L, dt, nt, ns = 10.0, 0.005, 201, 5
    every = nt // ns  # the reference cadence: freq = nt / num_snapshots
    b = FusedBatch(ns_list, L, dt, krylov_dim=10, mu2=1.0, snapshot_every=every,
                   max_snapshots=ns, cells_per_thread=cpt, pass_kernel=kernel)
    b.run(0)  # settles the tile
    assert b.tile()["pass_kind"] == kernel
    cases = [make_case(n, 2, t, ring=(t == 1)) for t, n in enumerate(ns_list)]
    for t, (m, c, u0, v0) in enumerate(cases):
        b.set_initial(t, m, c, u0, v0)
    b.run(nt - 1)
    for t, n in enumerate(ns_list):
        m, c, u0, v0 = cases[t]
        ur, vr = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=10, mu2=1.0)
        u, v = b.snapshots(t)
        assert u.shape[0] == ns
        for k in range(ns):
            assert rel_l2(u[k], ur[k]) < TOL, (t, k)
        for k in range(1, ns):
            assert rel_l2(v[k], vr[k]) < 1e-3, (t, k)  # first-order velocity, FP32 Krylov


@pytest.mark.parametrize("method", ["gautschi", "sewi"])
def test_fused_kge_3d_vs_oracle(orc, method):
    from nls.batch import FusedBatch
    from nls.samplers import bubbles_3d
    n, L, dt, nt, ns = 20, 8.0, 0.004, 61, 3
    for q6 in (True, False):
        u0, v0 = bubbles_3d(n, L, 3, 1)
        m = np.ones((n, n, n))
        c = np.ones((n, n, n))
        b = FusedBatch([n], L, dt, method=method, dim=3, krylov_dim=12, mu2=1.0, q6=q6,
                       snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=8)
        b.set_initial(0, m, c, u0, v0)
        b.run(nt - 1)
        u, _ = b.snapshots(0)
        ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=12, mu2=1.0, dim=3,
                            q6=q6, method={"gautschi": 0, "sewi": 2}[method])
        for k in range(ns):
            assert rel_l2(u[k], ur[k]) < TOL, (q6, k)


def make_case_3d(n, seed, traj, L=8.0):
    from nls.samplers import bubbles_3d
    rng = np.random.default_rng(seed * 100 + traj)
    u0, v0 = bubbles_3d(n, L, seed, traj)
    m = 0.8 + 0.4 * rng.random((n, n, n))
    c = 0.8 + 0.4 * rng.random((n, n, n))
    return m, c, u0, v0


_ORACLE_3D = {}


# direct 7-point stencil, plane march with 1024- and 2048-cell in-plane ranges; the
# sizes give one range (20^3), ranges that start mid-row (50^3: 896 cells) and several
# plane segments (50^3: 3 of 17 planes; 70^3: 4 of 18)
@pytest.mark.parametrize("kernel,cpt", [(0, 8), (4, 4), (4, 8)])
@pytest.mark.parametrize("q6", [True, False])
def test_fused_kge_3d_uneven_vs_oracle(orc, kernel, cpt, q6):
    from nls.batch import FusedBatch
    ns_list = [20, 50, 70]
    L, dt, nt, ns = 8.0, 0.004, 41, 2
    b = FusedBatch(ns_list, L, dt, method="sewi", dim=3, krylov_dim=10, mu2=1.0, q6=q6,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=cpt,
                   pass_kernel=kernel)
    cases = [make_case_3d(n, 3, t) for t, n in enumerate(ns_list)]
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(nt - 1)
    assert b.tile()["pass_kind"] == kernel
    for t, n in enumerate(ns_list):
        key = (n, q6)
        if key not in _ORACLE_3D:  # the oracle run does not depend on kernel / tile
            m, c, u0, v0 = cases[t]
            _ORACLE_3D[key] = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=10,
                                          mu2=1.0, dim=3, q6=q6, method=2)[0]
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], _ORACLE_3D[key][1]) < TOL, (t, n)


def test_fused_3d_plane_march_alone_equals_in_batch_bitwise():
    from nls.batch import FusedBatch
    mix = [24, 70, 17, 40]
    target = 1
    cases = [make_case_3d(n, 5, t) for t, n in enumerate(mix)]
    out = []
    for sizes, idx in ((mix, target), ([mix[target]], 0)):
        b = FusedBatch(sizes, 8.0, 0.004, dim=3, krylov_dim=10, mu2=1.0, cells_per_thread=4,
                       pass_kernel=4)
        for t, n in enumerate(sizes):
            b.set_initial(t, *(cases[t] if len(sizes) > 1 else cases[target]))
        b.run(20)
        out.append(b.state(idx)[0])
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("cpt,kernel", [(8, 1), (32, 2), (16, 3)])
def test_fused_alone_equals_in_batch_bitwise(cpt, kernel):
    from nls.batch import FusedBatch
    L, dt, steps = 10.0, 0.005, 40
    mix = [64, 40, 96, 64, 300]
    target = 2
    cases = [make_case(n, 4, t) for t, n in enumerate(mix)]
    b = FusedBatch(mix, L, dt, krylov_dim=8, mu2=1.0, cells_per_thread=cpt, pass_kernel=kernel)
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(steps)
    u_mix, _ = b.state(target)
    a = FusedBatch([mix[target]], L, dt, krylov_dim=8, mu2=1.0, cells_per_thread=cpt,
                   pass_kernel=kernel)
    a.set_initial(0, *cases[target])
    a.run(steps)
    u_alone, _ = a.state(0)
    assert np.array_equal(u_mix, u_alone)


@pytest.mark.parametrize("m", [3, 4, 5])
def test_fused_basis_only_pass_small_krylov(orc, m):
    """The basis-only pass kernel at the smallest Krylov dimensions, where the last
    pass is the j = 1 or j = 2 variant and stores w' for the combine."""
    from nls.batch import FusedBatch
    ns_list = [64, 40]
    L, dt, nt = 10.0, 0.005, 41
    cases = [make_case(n, 7, t) for t, n in enumerate(ns_list)]
    b = FusedBatch(ns_list, L, dt, krylov_dim=m, mu2=1.0, snapshot_every=nt // 2,
                   max_snapshots=2, cells_per_thread=16, pass_kernel=3)
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(nt - 1)
    assert b.tile()["pass_kind"] == 3
    for t, n in enumerate(ns_list):
        m_, c, u0, v0 = cases[t]
        ur, _ = orc.kge_run(n, L, dt * nt, nt, 2, m_, c, u0, v0, krylov=m, mu2=1.0)
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], ur[1]) < TOL, t


def test_fused_graph_replay_equals_eager():
    from nls.batch import FusedBatch
    L, dt, steps = 10.0, 0.005, 30
    ns_list = [50, 70]
    cases = [make_case(n, 5, t) for t, n in enumerate(ns_list)]
    out = []
    for graph in (True, False):
        b = FusedBatch(ns_list, L, dt, krylov_dim=9, mu2=1.0, use_graph=graph, cells_per_thread=16)
        for t, cs in enumerate(cases):
            b.set_initial(t, *cs)
        for _ in range(3):  # several replays with host synchronisation between them
            b.run(steps // 3)
        out.append([b.state(t)[0] for t in range(len(ns_list))])
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_fused_autotuned_tile_matches_oracle(orc):
    from nls.batch import FusedBatch
    n, L, dt, nt = 80, 10.0, 0.005, 41
    m, c, u0, v0 = make_case(n, 6, 0)
    b = FusedBatch([n], L, dt, krylov_dim=0, mu2=1.0, snapshot_every=nt // 2, max_snapshots=2)
    b.set_initial(0, m, c, u0, v0)
    b.run(nt - 1)
    info = b.info()
    assert 8 <= info["krylov_dim"] <= 16
    u, _ = b.snapshots(0)
    ur, _ = orc.kge_run(n, L, dt * nt, nt, 2, m, c, u0, v0, krylov=info["krylov_dim"], mu2=1.0)
    assert rel_l2(u[1], ur[1]) < TOL


def test_fused_zero_field_stays_zero():
    """Q14 edge: a zero field (zero Krylov seed) must give 0, not NaN."""
    from nls.batch import FusedBatch
    n = 32
    b = FusedBatch([n], 10.0, 0.005, krylov_dim=8, cells_per_thread=8)
    z = np.zeros((n, n))
    b.set_initial(0, np.ones((n, n)), np.ones((n, n)), z, z)
    b.run(5)
    u, _ = b.state(0)
    assert np.all(u == 0.0)


@pytest.mark.parametrize("kernel,cpt", [(0, 8), (2, 8), (2, 24)])  # direct / row march
def test_fused_nlse_vs_oracle(orc, kernel, cpt):
    """Config-2 shape at test size: complex64 fused Strang vs the complex128
    oracle (reference Lanczos variant), rel-L2 and mass."""
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    ns_list = [48, 40, 130]
    L, dt, nt, ns = 10.0, 0.005, 101, 5
    b = FusedBatch(ns_list, L, dt, equation="nlse", method="strang", krylov_dim=12,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=cpt,
                   pass_kernel=kernel)
    psis = []
    for t, n in enumerate(ns_list):
        psi, _ = nlse_packets_2d(n, L, 1, t)
        psis.append(psi)
        b.set_initial(t, np.ones((n, n)), np.ones((n, n)), psi)
    b.run(nt - 1)
    assert b.tile()["pass_kind"] == kernel
    for t, n in enumerate(ns_list):
        out, _ = b.snapshots(t)
        ref = orc.nlse_run(n, L, dt * nt, nt, ns, np.ones((n, n)), np.ones((n, n)), psis[t],
                           krylov=12, method=0, variant=0)
        for k in range(ns):
            assert rel_l2(out[k], ref[k]) < 1e-4, (t, k)
        m0 = np.sum(np.abs(ref[0]) ** 2)
        assert abs(np.sum(np.abs(out[-1]) ** 2) - np.sum(np.abs(ref[-1]) ** 2)) / m0 < 1e-5



@pytest.mark.parametrize("amp,dt,k", [(0.9, 0.005, 50), (12.0, 0.01, 5)])
@pytest.mark.parametrize("variant", ["reference", "full"])
def test_fused_nlse_constant_state_closed_form(amp, dt, k, variant):
    """A constant field is an eigenvector of the Neumann Laplacian with eigenvalue 0, so the
    Krylov factor is the identity and each Strang step multiplies psi by
    exp(i dt m |psi|^2): after k steps psi = A exp(i k dt m |A|^2).  The half-step phase
    0.5 dt m |A|^2 is 2e-3 for |A| = 0.9 (small-phase series) and 0.72 for |A| = 12 at
    dt = 0.01 (the library sincos).  The complex64 field's rounding moves |psi|^2 by about
    1e-7 per step, which the phase of 1.44 rad per step at |A| = 12 turns into 3.6e-5 after
    50 steps: that case runs 5 steps."""
    from nls.batch import FusedBatch
    n, mval = 40, 1.0
    A = amp * np.exp(0.7j)
    for kernel, cpt in [(2, 8), (0, 8)]:
        b = FusedBatch([n], 10.0, dt, equation="nlse", method="strang", krylov_dim=8,
                       cells_per_thread=cpt, pass_kernel=kernel, variant=variant)
        b.set_initial(0, np.full((n, n), mval), np.ones((n, n)),
                      np.full((n, n), A, dtype=np.complex128))
        b.run(k)
        psi, _ = b.state(0)
        a32 = complex(np.complex64(A))  # the field is held in complex64
        exact = a32 * np.exp(1j * k * dt * mval * abs(a32) ** 2)
        err = np.max(np.abs(psi - exact)) / abs(A)
        b.close()
        assert err < 1e-5, (kernel, err)


@pytest.mark.parametrize("krylov,stalls", [(3, 2), (5, 2), (8, 0), (10, 0)])
def test_fused_krylov_monitor_vs_oracle(orc, krylov, stalls):
    """Krylov monitor after one step equals the last-vector weight |y_{m-1}| / |y| of the
    oracle's reference-variant Lanczos (T[m-1][m-1] = 0) on the same inputs: cos(dt W) u
    and sinc^2(dt W) (-m u^3), the larger of the two; steps above 1e-6 are counted."""
    from nls.batch import FusedBatch
    n, L, dt = 64, 10.0, 0.005
    m, c, u0, v0 = make_case(n, 8, 0)
    b = FusedBatch([n], L, dt, krylov_dim=krylov, mu2=1.0, cells_per_thread=16)
    b.set_initial(0, m, c, u0, v0)
    b.run(1)
    mon = b.monitor(0)
    dx = 2 * L / (n - 1)
    op = orc.laplacian_2d(n, dx, dx, c, negate=True, shift=1.0)
    tails = []
    for x, kind in ((u0, 0), (-m * u0 ** 3, 1)):  # F_COS_SQRT, F_SINC2_SQRT
        beta, _, T = orc.lanczos_real(op, x.ravel(), krylov, 0)
        y = orc.small_matfunc_real(T, kind, dt) * beta
        tails.append(abs(y[-1]) / np.linalg.norm(y))
    assert abs(mon["max_tail"] - max(tails)) < 1e-3 * max(tails), (mon, tails)
    assert mon["stalls"] == stalls == sum(t > 1e-6 for t in tails)
    b.run(4)  # counts accumulate per step
    assert b.monitor(0)["stalls"] == 5 * stalls


# full reorthogonalisation (classical Gram-Schmidt, FP64 coefficients) in every iteration,
# the reference's own Lanczos form (oracle variant 0)
@pytest.mark.parametrize("kernel,cpt", [(0, 8), (2, 24)])
def test_fused_reorth_always_vs_oracle(orc, kernel, cpt):
    from nls.batch import FusedBatch
    ns_list = [48, 64, 33, 130]
    L, dt, nt, ns = 10.0, 0.005, 101, 2
    b = FusedBatch(ns_list, L, dt, krylov_dim=10, mu2=1.0, snapshot_every=nt // ns,
                   max_snapshots=ns, cells_per_thread=cpt, pass_kernel=kernel, reorth="always")
    cases = [make_case(n, 9, t, ring=(t == 1)) for t, n in enumerate(ns_list)]
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(nt - 1)
    for t, n in enumerate(ns_list):
        m, c, u0, v0 = cases[t]
        ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=10, mu2=1.0)
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], ur[1]) < TOL, t
        mon = b.monitor(t)
        assert mon["reorth_on"] and mon["max_orth"] < 1e-5, mon


def test_fused_reorth_always_3d_plane_march_vs_oracle(orc):
    from nls.batch import FusedBatch
    ns_list = [20, 33]
    L, dt, nt, ns = 8.0, 0.004, 41, 2
    b = FusedBatch(ns_list, L, dt, method="sewi", dim=3, krylov_dim=10, mu2=1.0,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=4, pass_kernel=4,
                   reorth="always")
    cases = [make_case_3d(n, 4, t) for t, n in enumerate(ns_list)]
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(nt - 1)
    for t, n in enumerate(ns_list):
        m, c, u0, v0 = cases[t]
        ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=10, mu2=1.0, dim=3,
                            method=2)
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], ur[1]) < TOL, t


def test_fused_reorth_auto_restores_orthogonality(orc):
    """A 6 x 6 grid with m = 24 exhausts the spectrum, so the FP32 basis loses
    orthogonality (|<v_22, v_0>| ~ 1e-2).  Mode 2 measures it in every step and, at the
    next check (every 16 steps), starts reorthogonalising; mode 1 keeps it below 1e-5 from
    the start; with a threshold never reached the loss is only measured."""
    from nls.batch import FusedBatch
    n, L, dt, steps = 6, 10.0, 0.005, 24
    m, c, u0, v0 = make_case(n, 8, 0)
    res = {}
    for name, mode, tol in (("measured", "auto", 1e30), ("auto", "auto", 0.0),
                            ("always", "always", 0.0)):
        b = FusedBatch([n], L, dt, krylov_dim=24, mu2=1.0, cells_per_thread=4, pass_kernel=0,
                       reorth=mode, reorth_tol=tol)
        b.set_initial(0, m, c, u0, v0)
        b.run(steps)
        res[name] = (b.monitor(0), b.state(0)[0])
    assert not res["measured"][0]["reorth_on"] and res["measured"][0]["max_orth"] > 1e-3
    assert res["auto"][0]["reorth_on"]
    assert res["always"][0]["reorth_on"] and res["always"][0]["max_orth"] < 1e-5
    # state after `steps` steps against the oracle's snapshot 1 of a 2 * steps run
    ur, _ = orc.kge_run(n, L, dt * 2 * steps, 2 * steps, 2, m, c, u0, v0, krylov=24, mu2=1.0)
    err = {k: rel_l2(v[1], ur[1]) for k, v in res.items()}
    print("rel-L2 vs oracle:", err)
    assert all(e < TOL for e in err.values()), err


def test_fused_reorth_inactive_is_bitwise_off():
    """Mode 2 below its threshold measures orthogonality but changes nothing."""
    from nls.batch import FusedBatch
    ns_list = [64, 40]
    cases = [make_case(n, 3, t) for t, n in enumerate(ns_list)]
    out = []
    for mode, tol in ((0, 0.0), ("auto", 1e30)):
        b = FusedBatch(ns_list, 10.0, 0.005, krylov_dim=8, mu2=1.0, cells_per_thread=24,
                       pass_kernel=2, lanes=1, reorth=mode, reorth_tol=tol)
        for t, cs in enumerate(cases):
            b.set_initial(t, *cs)
        b.run(15)
        out.append([b.state(t)[0] for t in range(len(ns_list))])
    for x, y in zip(*out):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("kernel,cpt", [(0, 8), (2, 16)])
def test_fused_nlse_reorth_always_vs_oracle(orc, kernel, cpt):
    """Complex Gram-Schmidt in every iteration (the reference's NLSE Lanczos form)."""
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    ns_list = [48, 40, 130]
    L, dt, nt, ns = 10.0, 0.005, 41, 2
    b = FusedBatch(ns_list, L, dt, equation="nlse", method="strang", krylov_dim=12,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=cpt,
                   pass_kernel=kernel, reorth="always")
    psis = []
    for t, n in enumerate(ns_list):
        psi, _ = nlse_packets_2d(n, L, 1, t)
        psis.append(psi)
        b.set_initial(t, np.ones((n, n)), np.ones((n, n)), psi)
    b.run(nt - 1)
    for t, n in enumerate(ns_list):
        out, _ = b.snapshots(t)
        ref = orc.nlse_run(n, L, dt * nt, nt, ns, np.ones((n, n)), np.ones((n, n)), psis[t],
                           krylov=12, method=0, variant=0)
        assert rel_l2(out[1], ref[1]) < 1e-4, t
        mon = b.monitor(t)
        assert mon["reorth_on"] and mon["max_orth"] < 1e-5, mon


def test_fused_nlse_reorth_auto_restores_orthogonality(orc):
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    n, L, dt, steps = 6, 10.0, 0.005, 24
    psi, _ = nlse_packets_2d(n, L, 1, 0)
    res = {}
    for name, mode, tol in (("measured", "auto", 1e30), ("auto", "auto", 0.0),
                            ("always", "always", 0.0)):
        b = FusedBatch([n], L, dt, equation="nlse", method="strang", krylov_dim=24,
                       cells_per_thread=4, pass_kernel=0, reorth=mode, reorth_tol=tol)
        b.set_initial(0, np.ones((n, n)), np.ones((n, n)), psi)
        b.run(steps)
        res[name] = b.monitor(0)
    print(res)
    assert not res["measured"]["reorth_on"] and res["measured"]["max_orth"] > 1e-3
    assert res["auto"]["reorth_on"]
    assert res["always"]["reorth_on"] and res["always"]["max_orth"] < 1e-5


def test_fused_long_horizon_vs_oracle(orc):
    """3000 steps (a fifth of the uneven-batch configuration's nt = 15000) on a 128^2 ring:
    the FP32-Krylov / FP64-state error grows about linearly (7.8e-7 at the step-2500
    snapshot when recorded) and stays below the 1e-5 bar at every snapshot."""
    from nls.batch import FusedBatch
    n, L, dt, nt, ns = 128, 10.0, 0.005, 3001, 6
    m, c, u0, v0 = make_case(n, 2, 1, ring=True)
    b = FusedBatch([n], L, dt, krylov_dim=8, mu2=1.0, snapshot_every=nt // ns, max_snapshots=ns,
                   cells_per_thread=16, pass_kernel=2)
    b.set_initial(0, m, c, u0, v0)
    b.run(nt - 1)
    u, _ = b.snapshots(0)
    ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=8, mu2=1.0)
    errs = [rel_l2(u[k], ur[k]) for k in range(1, ns)]
    assert max(errs) < TOL, errs
    assert errs[-1] < 2e-6, errs


@pytest.mark.timeout(400)
def test_fused_kge_uneven_workload_full_horizon_vs_oracle(orc):
    """The uneven-batch workload's whole horizon on one of its grids: 128^2 ring/kink of
    seed 2, side 20, mu2 = 1, dt = 0.005, nt = 15000, a snapshot every 150 steps (100
    snapshots), the full-T Lanczos variant at the Krylov dimension the autotuner picks
    (m = 8), against the oracle's full-T Lanczos: the workload's own uniform m = c = 1
    and a varying m, c, both at the 1e-5 bar at every snapshot.  (The reference variant,
    T[m-1][m-1] = 0, is ill-conditioned on the uniform case at m = 8:
    test_reference_semantics_conditioning_with_uniform_fields.)"""
    from nls.batch import FusedBatch
    from nls.samplers import kink_ring_2d
    n, L, dt, nt, ns = 128, 10.0, 0.005, 15000, 100
    u0, v0, _ = kink_ring_2d(n, L, 2, 0)
    rng = np.random.default_rng(201)
    one = np.ones((n, n))
    fields = [(one, one), (0.8 + 0.4 * rng.random((n, n)), 0.8 + 0.4 * rng.random((n, n)))]
    b = FusedBatch([n, n], L, dt, krylov_dim=0, mu2=1.0, snapshot_every=nt // ns,
                   max_snapshots=ns, reorth="auto", variant="full")
    for t, (m, c) in enumerate(fields):
        b.set_initial(t, m, c, u0, v0)
    b.run(nt - 1)
    krylov = b.info()["krylov_dim"]
    assert krylov == 8, krylov
    ur, _ = orc.kge_run_batch(n, L, dt * nt, nt, ns, [f[0] for f in fields],
                              [f[1] for f in fields], [u0, u0], [v0, v0], krylov=krylov,
                              mu2=1.0, variant=1)
    for t, name in enumerate(("uniform", "varying")):
        u, _ = b.snapshots(t)
        errs = [rel_l2(u[k], ur[t][k]) for k in range(1, ns)]
        print(f"kge 128^2 x {nt} ({name} m, c; full T, m = {krylov}): rel-L2 at snapshots "
              f"25/50/75/99: {errs[24]:.2e} {errs[49]:.2e} {errs[74]:.2e} {errs[-1]:.2e}, "
              f"max {max(errs):.2e}")
        assert max(errs) < TOL, (name, max(errs))


@pytest.mark.timeout(300)
def test_fused_kge_headline_trajectory_full_variant_vs_oracle(orc):
    """A trajectory of the headline workload (300^2, side 20, uniform m = c = 1, mu2 = 1,
    dt = 0.005, ring/kink of seed 2) over 1500 steps with the full-T variant at the
    autotuned m, against the oracle's full-T Lanczos: 1e-5 at every one of 10 snapshots."""
    from nls.batch import FusedBatch
    from nls.samplers import kink_ring_2d
    n, L, dt, nt, ns = 300, 10.0, 0.005, 1500, 10
    u0, v0, _ = kink_ring_2d(n, L, 2, 3)
    one = np.ones((n, n))
    b = FusedBatch([n], L, dt, krylov_dim=0, mu2=1.0, snapshot_every=nt // ns,
                   max_snapshots=ns, reorth="auto", variant="full")
    b.set_initial(0, one, one, u0, v0)
    b.run(nt - 1)
    krylov = b.info()["krylov_dim"]
    u, _ = b.snapshots(0)
    ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, one, one, u0, v0, krylov=krylov, mu2=1.0,
                        variant=1)
    errs = [rel_l2(u[k], ur[k]) for k in range(1, ns)]
    print(f"kge 300^2 x {nt} (full T, m = {krylov}): rel-L2 {['%.2e' % e for e in errs]}")
    assert max(errs) < TOL, errs


def _oracle_at_step(orc, n, L, dt, steps, case, krylov, variant, **kw):
    """The oracle's state after `steps` steps (its snapshot at step `steps` of a run of
    2 * steps steps with two snapshots)."""
    m, c, u0, v0 = case
    ur, _ = orc.kge_run(n, L, dt * 2 * steps, 2 * steps, 2, m, c, u0, v0, krylov=krylov,
                        variant=variant, **kw)
    return ur[1]


# (pass kernel, tile): direct stencil, row march
@pytest.mark.parametrize("kernel,cpt", [(0, 8), (2, 16)])
@pytest.mark.parametrize("krylov", [4, 6])
def test_fused_full_variant_vs_oracle(orc, kernel, cpt, krylov):
    """The full-T variant's extra pass (alpha_{m-1}) against the oracle's full-T Lanczos on
    small Krylov spaces, where T[m-1][m-1] matters: the two oracle variants differ by
    ~1e-3 (m = 4) and ~2.5e-4 (m = 6) after 20 steps of dt = 0.01 (checked here), so a
    missing, misplaced or mis-scaled alpha_{m-1} fails the 1e-6 bar by orders of
    magnitude; uneven batch, varying m and c."""
    from nls.batch import FusedBatch
    ns_list = [32, 48, 17]
    L, dt, steps = 10.0, 0.01, 20
    cases = [make_case(n, 4, t) for t, n in enumerate(ns_list)]
    b = FusedBatch(ns_list, L, dt, krylov_dim=krylov, mu2=1.0, snapshot_every=steps,
                   max_snapshots=2, cells_per_thread=cpt, pass_kernel=kernel, variant="full")
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(steps)
    assert b.tile()["pass_kind"] == kernel
    for t, n in enumerate(ns_list):
        r1 = _oracle_at_step(orc, n, L, dt, steps, cases[t], krylov, 1, mu2=1.0)
        r0 = _oracle_at_step(orc, n, L, dt, steps, cases[t], krylov, 0, mu2=1.0)
        assert rel_l2(r0, r1) > 1e-5, (t, rel_l2(r0, r1))  # the case tells the variants apart
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], r1) < 1e-6, (t, n, rel_l2(u[1], r1))


@pytest.mark.parametrize("kernel,cpt", [(0, 8), (4, 4), (4, 8)])
def test_fused_full_variant_3d_vs_oracle(orc, kernel, cpt):
    """The full-T variant on the 3D passes (direct 7-point stencil, plane march with 1024-
    and 2048-cell ranges), sEWI, m = 5, against the oracle's full-T Lanczos."""
    from nls.batch import FusedBatch
    ns_list = [20, 33]
    L, dt, steps, krylov = 8.0, 0.02, 10, 5
    cases = [make_case_3d(n, 3, t) for t, n in enumerate(ns_list)]
    b = FusedBatch(ns_list, L, dt, method="sewi", dim=3, krylov_dim=krylov, mu2=1.0,
                   snapshot_every=steps, max_snapshots=2, cells_per_thread=cpt,
                   pass_kernel=kernel, variant="full")
    for t, cs in enumerate(cases):
        b.set_initial(t, *cs)
    b.run(steps)
    assert b.tile()["pass_kind"] == kernel
    for t, n in enumerate(ns_list):
        kw = dict(mu2=1.0, dim=3, method=2)
        r1 = _oracle_at_step(orc, n, L, dt, steps, cases[t], krylov, 1, **kw)
        r0 = _oracle_at_step(orc, n, L, dt, steps, cases[t], krylov, 0, **kw)
        assert rel_l2(r0, r1) > 1e-5, (t, rel_l2(r0, r1))
        u, _ = b.snapshots(t)
        assert rel_l2(u[1], r1) < 1e-6, (t, n, rel_l2(u[1], r1))


@pytest.mark.parametrize("dim,kernel,cpt", [(2, 0, 8), (2, 2, 8), (3, 4, 4), (3, 0, 8)])
def test_fused_nlse_full_variant_vs_oracle(orc, dim, kernel, cpt):
    """The complex64 NLSE passes with the full-T variant against the oracle's full-T
    complex Lanczos (Strang, m = 3, dt = 0.05, where the oracle's two variants differ by
    7e-2 in 2D and 1e-3 in 3D after 10 steps)."""
    from nls.batch import FusedBatch
    L, dt, steps, krylov = 10.0, 0.05, 10, 3
    if dim == 2:
        from nls.samplers import nlse_packets_2d
        n = 40
        psi0, _ = nlse_packets_2d(n, L, 1, 2)
        rng = np.random.default_rng(7)
        m, c = 0.8 + 0.4 * rng.random((n, n)), 0.8 + 0.4 * rng.random((n, n))
    else:
        n = 20
        m, c, psi0 = nlse_case_3d(n, L, 1)
    b = FusedBatch([n], L, dt, equation="nlse", method="strang", dim=dim, krylov_dim=krylov,
                   snapshot_every=steps, max_snapshots=2, cells_per_thread=cpt,
                   pass_kernel=kernel, variant="full")
    b.set_initial(0, m, c, psi0)
    b.run(steps)
    assert b.tile()["pass_kind"] == kernel
    out, _ = b.snapshots(0)
    ref = {v: orc.nlse_run(n, L, dt * 2 * steps, 2 * steps, 2, m, c, psi0, krylov=krylov,
                           method=0, variant=v, dim=dim)[1] for v in (0, 1)}
    assert rel_l2(ref[0], ref[1]) > 1e-5, rel_l2(ref[0], ref[1])
    assert rel_l2(out[1], ref[1]) < 1e-5, rel_l2(out[1], ref[1])


def test_fused_nlse_full_horizon_vs_oracle(orc):
    """The NLSE workload's own size and horizon (200^2 on [-10, 10]^2, m = 12,
    dt = 0.005, nt = 2000, packets of seed 1): complex64 fused Strang against the
    complex128 oracle at 5 snapshots: rel-L2 1.5e-6 at the last snapshot when recorded
    (bar 1e-5 at every snapshot, the KGE field bar), the mass equal to the oracle's (which
    carries the reference's own ghost-copy mass loss) to 2.4e-8 (bar 1e-6)."""
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    n, L, dt, nt, ns = 200, 10.0, 0.005, 2000, 5
    psi0, _ = nlse_packets_2d(n, L, 1, 0)
    one = np.ones((n, n))
    b = FusedBatch([n], L, dt, equation="nlse", method="strang", krylov_dim=12,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=8, pass_kernel=2)
    b.set_initial(0, one, one, psi0)
    b.run(nt - 1)
    out, _ = b.snapshots(0)
    ref = orc.nlse_run(n, L, dt * nt, nt, ns, one, one, psi0, krylov=12, method=0, variant=0)
    errs = [rel_l2(out[k], ref[k]) for k in range(ns)]
    m0 = np.sum(np.abs(ref[0]) ** 2)
    dm = [abs(np.sum(np.abs(out[k]) ** 2) - np.sum(np.abs(ref[k]) ** 2)) / m0 for k in range(ns)]
    print(f"nlse 200^2 x {nt}: rel-L2 {['%.2e' % e for e in errs]}, mass gap {max(dm):.2e}")
    assert max(errs) < 1e-5, errs
    assert max(dm) < 1e-6, dm


def test_fused_kge_3d_long_horizon_vs_oracle(orc):
    """1000 sEWI steps of a 32^3 trajectory (m = 12, dt = 0.004, bubbles of seed 3, varying
    m and c): the FP32-Krylov / FP64-state error stays below the 1e-5 bar at every
    snapshot (2.2e-7 at the last one when recorded)."""
    from nls.batch import FusedBatch
    n, L, dt, nt, ns = 32, 8.0, 0.004, 1000, 5
    m, c, u0, v0 = make_case_3d(n, 3, 2)
    b = FusedBatch([n], L, dt, method="sewi", dim=3, krylov_dim=12, mu2=1.0,
                   snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=4, pass_kernel=4)
    b.set_initial(0, m, c, u0, v0)
    b.run(nt - 1)
    u, _ = b.snapshots(0)
    ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=12, mu2=1.0, dim=3,
                        method=2)
    errs = [rel_l2(u[k], ur[k]) for k in range(1, ns)]
    print(f"kge 32^3 x {nt}: rel-L2 {['%.2e' % e for e in errs]}")
    assert max(errs) < TOL, errs


@pytest.mark.timeout(400)
def test_fused_kge_config4_size_vs_oracle(orc):
    """Config 4 at its own size: one 220^3 trajectory on [-8, 8]^3 (7-point operator with
    the reference's cross-plane couplings), sEWI, m = 12, dt = 0.004, mu2 = 1, Gaussian
    bubbles of seed 3, on the plane march with 1024-cell ranges: 10 steps against the FP64
    oracle (full-T Lanczos in both) at the 1e-5 bar."""
    from nls.batch import FusedBatch
    from nls.samplers import bubbles_3d
    n, L, dt, steps, krylov = 220, 8.0, 0.004, 10, 12
    u0, v0 = bubbles_3d(n, L, 3, 0)
    one = np.ones((n, n, n))
    b = FusedBatch([n], L, dt, method="sewi", dim=3, krylov_dim=krylov, mu2=1.0,
                   snapshot_every=steps, max_snapshots=2, cells_per_thread=4, pass_kernel=4,
                   variant="full")
    b.set_initial(0, one, one, u0, v0)
    b.run(steps)
    assert b.tile()["pass_kind"] == 4
    u, _ = b.snapshots(0)
    b.close()
    # the oracle's snapshot every step (freq = nt / ns = 1); snapshot `steps` is the state
    ur, _ = orc.kge_run(n, L, dt * (steps + 1), steps + 1, steps + 1, one, one, u0, v0,
                        krylov=krylov, mu2=1.0, dim=3, method=2, variant=1)
    err = rel_l2(u[1], ur[steps])
    print(f"kge 220^3 x {steps} (full T, m = {krylov}): rel-L2 {err:.2e}")
    assert err < TOL, err


@pytest.mark.timeout(300)
def test_fused_kge_config4_batch_of_40_properties():
    """The 40-trajectory 220^3 batch resident in HBM (about 55 GB with m = 12): after 5
    sEWI steps every trajectory is finite and nonzero, and trajectory 7 is bitwise the same
    as when it runs alone with the same tile."""
    from nls.batch import FusedBatch
    n, L, dt, steps, krylov = 220, 8.0, 0.004, 5, 12
    kw = dict(method="sewi", dim=3, krylov_dim=krylov, mu2=1.0, cells_per_thread=4,
              pass_kernel=4, variant="full")
    b = FusedBatch([n] * 40, L, dt, **kw)
    for t in range(40):
        b.sample_initial(t, "bubbles_3d", 3, m=1.0, c=1.0)
    b.run(steps)
    assert b.info()["device_bytes"] > 40 * n ** 3 * 100
    states = {}
    for t in range(40):
        u, _ = b.state(t)
        assert np.all(np.isfinite(u)) and np.any(u != 0.0), t
        if t == 7:
            states[t] = u
    b.close()
    a = FusedBatch([n], L, dt, **kw)
    a.sample_initial(0, "bubbles_3d", 3, traj_index=7, m=1.0, c=1.0)
    a.run(steps)
    u, _ = a.state(0)
    a.close()
    assert np.array_equal(u, states[7])


@pytest.mark.parametrize("dim", [2, 3])
def test_fused_kge_energy_tracks_oracle(orc, dim):
    """Discrete energy E = 1/2 |v|^2 + 1/2 u.A u + 1/4 sum m u^4 (A = -L_c + mu2,
    compare_utils_real_2d.py:112-131 plus the mass term) of the fused run against the
    FP64 oracle's at every snapshot: both carry the same O(tau) oscillation of the
    first-order velocity, so their difference is the fused path's own error: 1.4e-8 (2D)
    and 7e-9 (3D) of E when recorded, bar 1e-6.  The fused drift itself (about 6e-3 here,
    the O(tau) oscillation) must equal the oracle's to the same bar."""
    from nls.batch import FusedBatch
    if dim == 2:
        n, L, dt, nt, ns, krylov = 96, 10.0, 0.005, 801, 9, 8
        m, c, u0, v0 = make_case(n, 2, 1, ring=True)
        op = orc.laplacian_2d(n, 2 * L / (n - 1), 2 * L / (n - 1), c.ravel(), negate=True,
                              shift=1.0)
    else:
        n, L, dt, nt, ns, krylov = 24, 8.0, 0.004, 401, 5, 12
        m, c, u0, v0 = make_case_3d(n, 3, 0)
        op = orc.laplacian_3d(n, 2 * L / (n - 1), c.ravel(), negate=True, shift=1.0)
    b = FusedBatch([n], L, dt, dim=dim, krylov_dim=krylov, mu2=1.0, snapshot_every=nt // ns,
                   max_snapshots=ns, cells_per_thread=8)
    b.set_initial(0, m, c, u0, v0)
    b.run(nt - 1)
    u, v = b.snapshots(0)
    ur, vr = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=krylov, mu2=1.0, dim=dim)

    def energy(uu, vv):
        uu, vv = uu.ravel(), vv.ravel()
        return 0.5 * vv @ vv + 0.5 * uu @ op.matvec(uu) + 0.25 * np.sum(m.ravel() * uu ** 4)
    Ef = np.array([energy(u[k], v[k]) for k in range(1, ns)])
    Eo = np.array([energy(ur[k], vr[k]) for k in range(1, ns)])
    gap = np.abs(Ef - Eo).max() / abs(Eo[0])
    drift_f = np.abs(Ef - Ef[0]).max() / abs(Ef[0])
    drift_o = np.abs(Eo - Eo[0]).max() / abs(Eo[0])
    print(f"dim {dim}: energy gap {gap:.2e}, drift fused {drift_f:.3e}, oracle {drift_o:.3e}")
    assert gap < 1e-6, (gap, drift_f, drift_o)
    assert abs(drift_f - drift_o) < 1e-6, (drift_f, drift_o)


def test_fused_reorth_auto_acts_per_trajectory():
    """Mode 2 in a batch of several lanes: only the trajectory whose basis lost
    orthogonality (6 x 6 grid, m = 24) switches on; the graphs are recaptured with the
    reorthogonalisation kernels mid-run, and the other trajectories' results stay bitwise
    those of a run without reorthogonalisation."""
    from nls.batch import FusedBatch
    sizes = [40, 6, 64, 33]
    cases = [make_case(n, 8, t) for t, n in enumerate(sizes)]
    out = {}
    for mode in (0, "auto"):
        b = FusedBatch(sizes, 10.0, 0.005, krylov_dim=24, mu2=1.0, cells_per_thread=8,
                       pass_kernel=0, lanes=2, reorth=mode, snapshot_every=10, max_snapshots=5)
        for t, cs in enumerate(cases):
            b.set_initial(t, *cs)
        b.run(20)
        b.run(20)
        out[mode] = ([b.state(t)[0] for t in range(len(sizes))],
                     [b.monitor(t) for t in range(len(sizes))],
                     [b.snapshots(t)[0] for t in range(len(sizes))])
    mons = out["auto"][1]
    assert [m["reorth_on"] for m in mons] == [False, True, False, False], mons
    for t in (0, 2, 3):
        assert np.array_equal(out[0][0][t], out["auto"][0][t]), t
        assert np.array_equal(out[0][2][t], out["auto"][2][t]), t


def nlse_case_3d(n, L, t):
    x = np.linspace(-L, L, n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rng = np.random.default_rng(40 + t)
    cx, cy, cz = rng.uniform(-2, 2, 3)
    psi = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / 4) * np.exp(1j * (0.7 * X - 0.3 * Z))
    m = 0.8 + 0.4 * rng.random((n, n, n))
    c = 0.8 + 0.4 * rng.random((n, n, n))
    return m, c, psi


# NLSE in 3D: direct stencil and plane march (one range at 20^3, ranges starting mid-row
# and 4 plane segments at 70^3), both operator variants
@pytest.mark.parametrize("kernel,cpt", [(0, 8), (4, 4)])
@pytest.mark.parametrize("q6", [True, False])
def test_fused_nlse_3d_vs_oracle(orc, kernel, cpt, q6):
    from nls.batch import FusedBatch
    ns_list = [20, 70]
    L, dt, nt, ns = 8.0, 0.002, 21, 2
    b = FusedBatch(ns_list, L, dt, equation="nlse", method="strang", dim=3, krylov_dim=12,
                   q6=q6, snapshot_every=nt // ns, max_snapshots=ns, cells_per_thread=cpt,
                   pass_kernel=kernel)
    cases = [nlse_case_3d(n, L, t) for t, n in enumerate(ns_list)]
    for t, (m, c, psi) in enumerate(cases):
        b.set_initial(t, m, c, psi)
    b.run(nt - 1)
    assert b.tile()["pass_kind"] == kernel
    for t, n in enumerate(ns_list):
        key = ("nlse", n, q6)
        if key not in _ORACLE_3D:  # independent of kernel and tile
            m, c, psi = cases[t]
            _ORACLE_3D[key] = orc.nlse_run(n, L, dt * nt, nt, ns, m, c, psi, krylov=12, dim=3,
                                           q6=q6)
        out, _ = b.snapshots(t)
        assert rel_l2(out[1], _ORACLE_3D[key][1]) < 1e-4, (t, n)


# uniform-field path of the row march (2D) and plane march (3D), KGE and NLSE
@pytest.mark.parametrize("equation,dim,kernel,cpt", [("kge", 2, 2, 16), ("kge", 3, 4, 4),
                                                     ("nlse", 2, 2, 8), ("nlse", 3, 4, 4)])
def test_fused_uniform_fields_equal_streamed_bitwise(equation, dim, kernel, cpt):
    """A batch whose trajectories all have spatially uniform m and c runs the kernel
    variants that take the values from the descriptor instead of streaming the fields:
    bitwise the result of streaming them (always_load_fields), and bitwise the result the
    same trajectory gets in a batch with a varying-field mate (which streams the fields;
    a fixed tile makes a trajectory's bits independent of its batch-mates)."""
    from nls.batch import FusedBatch
    from nls.samplers import nlse_packets_2d
    sizes = [40, 56] if dim == 2 else [20, 28]
    L = 10.0 if dim == 2 else 8.0

    def case(t):
        n = sizes[t]
        if equation == "kge":
            m, c, u0, v0 = make_case(n, 6, t) if dim == 2 else make_case_3d(n, 6, t)
        elif dim == 3:
            (m, c, u0), v0 = nlse_case_3d(n, L, t), None
        else:
            rng = np.random.default_rng(60 + t)
            m, c = 0.8 + 0.4 * rng.random((n, n)), 0.8 + 0.4 * rng.random((n, n))
            u0, v0 = nlse_packets_2d(n, L, 1, t)[0], None
        return m, c, u0, v0

    varying = [case(t) for t in range(2)]
    uniform = [(np.full(np.shape(m), mv), np.full(np.shape(c), cv), u0, v0)
               for (m, c, u0, v0), mv, cv in zip(varying, (0.7, 1.1), (1.3, 0.9))]

    def run(cases, load):
        kw = dict(equation="nlse", method="strang", krylov_dim=12) if equation == "nlse" else \
            dict(krylov_dim=8, mu2=1.0)
        b = FusedBatch(sizes, L, 0.004, dim=dim, cells_per_thread=cpt, pass_kernel=kernel,
                       always_load_fields=load, **kw)
        for t, (m, c, u0, v0) in enumerate(cases):
            if equation == "kge":
                b.set_initial(t, m, c, u0, v0)
            else:
                b.set_initial(t, m, c, u0)
        b.run(25)
        assert b.tile()["pass_kind"] == kernel
        return [b.state(t)[0] for t in range(len(sizes))]

    uni = run(uniform, False)       # uniform-field variants
    streamed = run(uniform, True)   # the same fields, streamed
    mixed = run([uniform[0], varying[1]], False)  # a varying mate: streamed variants
    for t in range(len(sizes)):
        assert np.array_equal(uni[t], streamed[t]), t
    assert np.array_equal(uni[0], mixed[0])
    assert not np.array_equal(uni[1], mixed[1])  # the mate's own fields took effect


def test_fused_uniform_fields_vs_oracle(orc):
    """Uniform m = 0.7, c = 1.3 (not 1) through the uniform-field path against the FP64
    oracle with the same constant fields."""
    from nls.batch import FusedBatch
    n, L, dt, nt, ns = 64, 10.0, 0.005, 201, 5
    _, _, u0, v0 = make_case(n, 2, 1, ring=True)
    m, c = np.full((n, n), 0.7), np.full((n, n), 1.3)
    b = FusedBatch([n], L, dt, krylov_dim=8, mu2=1.0, snapshot_every=nt // ns, max_snapshots=ns,
                   cells_per_thread=16, pass_kernel=2)
    b.set_initial(0, m, c, u0, v0)
    b.run(nt - 1)
    u, _ = b.snapshots(0)
    ur, _ = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, krylov=8, mu2=1.0)
    for k in range(ns):
        assert rel_l2(u[k], ur[k]) < TOL, k


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.mark.parametrize("dim,q6", [(2, True), (3, True), (3, False)])
def test_fused_kge_sv_bitwise_vs_oracle(orc, dim, q6):
    """The Stormer-Verlet batch (kg_single.cuh:129-179) runs the reference arithmetic in
    FP64: each operator row formed as the reference assembles it and summed in CSR column
    order without contraction.  With FP32-representable m and c (the batch keeps the
    fields in FP32) every snapshot of u and v is bitwise the oracle's, for an uneven batch
    with varying fields and for a trajectory with uniform ones (the uniform-field kernel)."""
    from nls.batch import FusedBatch
    if dim == 2:
        sizes, L, dt, nt, ns = [48, 33, 64], 10.0, 0.002, 201, 5
        cases = [make_case(n, 8, t) for t, n in enumerate(sizes)]
    else:
        sizes, L, dt, nt, ns = [20, 17], 8.0, 0.002, 61, 3
        cases = [make_case_3d(n, 8, t) for t, n in enumerate(sizes)]
    cases = [(_f32(m), _f32(c), u0, v0) for m, c, u0, v0 in cases]
    uniform = [(np.full_like(m, 0.75), np.full_like(c, 1.25), u0, v0) for m, c, u0, v0 in cases]
    for batch in (cases, uniform):
        b = FusedBatch(sizes, L, dt, method="sv", dim=dim, mu2=1.0, q6=q6,
                       snapshot_every=nt // ns, max_snapshots=ns)
        for t, cs in enumerate(batch):
            b.set_initial(t, *cs)
        b.run(nt - 1)
        for t, n in enumerate(sizes):
            m, c, u0, v0 = batch[t]
            ur, vr = orc.kge_run(n, L, dt * nt, nt, ns, m, c, u0, v0, method=1, mu2=1.0,
                                 dim=dim, q6=q6)
            u, v = b.snapshots(t)
            assert np.array_equal(u, ur), (t, np.abs(u - ur).max())
            assert np.array_equal(v, vr), (t, np.abs(v - vr).max())


def test_fused_pass_timing_counts_general_passes():
    """The profiling leg times the back-to-back general passes of each step with one event
    pair and reports their number: steps x (m - 3) passes, a positive average, and the
    algorithmic bytes of one pass."""
    from nls.batch import FusedBatch
    n = 96
    m, c, u0, v0 = make_case(n, 3, 0, ring=True)
    for kernel in (0, 2):
        b = FusedBatch([n, n], 10.0, 0.005, krylov_dim=8, mu2=1.0, cells_per_thread=8,
                       pass_kernel=kernel, variant="full")
        for t in range(2):
            b.set_initial(t, m, c, u0, v0)
        b.set_profiling(True)
        b.run(4)
        pt = b.pass_timing()
        assert pt["launches"] == 4 * (8 - 3), pt
        assert 0.0 < pt["avg_ms"] < 5.0, pt
        assert pt["bytes_per_launch"] > 0, pt
        b.close()
