"""CPU tests of the serving loop's host logic: the HBM-budget admission allocator."""


def test_pool_first_fit_and_coalescing():
    from nls.server import pool_simulate
    cap = 1 << 20
    r = pool_simulate(cap, [(0, 300_000), (0, 300_000), (0, 300_000), (0, 300_000),
                            (1, 1), (0, 200_000), (0, 200_000), (1, 0), (1, 2), (0, 700_000),
                            (1, 5), (0, 700_000)])
    a0, a1, a2, a3 = r[:4]
    assert a0 == 0 and a1 > a0 and a2 > a1
    assert a3 == -1                       # over budget: not admitted
    assert r[5] == a1                     # first fit reuses the freed hole
    assert r[6] == -1                     # the hole is smaller now; no room elsewhere
    assert r[8] == 200_192                # only block 5 left in use
    assert r[9] == -1                     # free space is split around block 5
    assert r[11] == 0                     # freeing it coalesces everything into one extent


def test_pool_rejects_everything_when_full():
    from nls.server import pool_simulate
    cap = 4096
    r = pool_simulate(cap, [(0, 4096), (0, 1), (1, 0), (0, 1)])
    assert r == [0, -1, 0, 0]
