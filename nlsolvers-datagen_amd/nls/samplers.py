"""Seeded initial-condition samplers for the benchmark configurations.

Shapes follow the reference's samplers; the parameter laws are the ones the
configurations state:
  gaussian_sum_2d   sum of Gaussians A exp(-|x-x0|^2/w^2)  (util.hpp:94-125 profile)
  kink_ring_2d      tanh((R - r)/w)                        (real_sampler.py:464-467)
  nlse_packets_2d   Gaussian packets A exp(-|x-x0|^2/w^2) exp(i(v.(x-x0) + phi))
                                                           (nlse_sampler.py:43-100 form)
  bubbles_3d        sum of 3D Gaussian bubbles A exp(-|x-x0|^2/w^2)  (util.hpp:94-125
                    profile; the configurations' 3D Gaussian bubbles)
Grids are linspace(-L, L, n) with indexing="ij" (real_sampler.py:17-25).

Every random parameter is a Philox4x32-10 draw keyed by the seed, with counter
(slot, trajectory, tag, 0), so a trajectory's initial data do not depend on how
many trajectories are drawn or in which order (the reference uses the global
legacy np.random stream, real_launcher_2d.py:390-395).
"""
import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Vectorised Philox4x32-10: ctr uint32 array (..., 4), key (k0, k1)."""
    c = np.array(ctr, dtype=np.uint32)
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0, k1 = np.uint32(key[0] & 0xFFFFFFFF), np.uint32(key[1] & 0xFFFFFFFF)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        n0 = ((p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)) & _MASK
        n1 = p1 & _MASK
        n2 = ((p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)) & _MASK
        n3 = p0 & _MASK
        c0, c1, c2, c3 = n0, n1, n2, n3
        k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


class ParamStream:
    """Uniform draws in (0,1) for one trajectory: slot k -> Philox((k, traj, tag, 0), seed)."""

    def __init__(self, seed, traj, tag=0x5EED):
        self.seed = int(seed)
        self.traj = int(traj) & 0xFFFFFFFF
        self.tag = int(tag) & 0xFFFFFFFF
        self.slot = 0

    def uniform(self, lo=0.0, hi=1.0, size=None):
        cnt = 1 if size is None else int(np.prod(size))
        groups = (cnt + 3) // 4
        ctr = np.zeros((groups, 4), np.uint32)
        ctr[:, 0] = np.arange(self.slot, self.slot + groups, dtype=np.uint64).astype(np.uint32)
        ctr[:, 1] = self.traj
        ctr[:, 2] = self.tag
        self.slot += groups
        x = philox4x32_10(ctr, (self.seed & 0xFFFFFFFF, self.seed >> 32)).ravel()[:cnt]
        u = (x.astype(np.float64) + 0.5) * 2.0 ** -32
        v = lo + (hi - lo) * u
        return float(v[0]) if size is None else v.reshape(size)

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + min(int(self.uniform() * (hi - lo + 1)), hi - lo)

    def normal(self, size=None):
        u1 = self.uniform(size=size)
        u2 = self.uniform(size=size)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)


def grid_2d(n, L):
    x = np.linspace(-L, L, n)
    return np.meshgrid(x, x, indexing="ij")


def gaussian_sum_2d(n, L, seed, traj, k=3, amp=(0.5, 2.0), width=(0.5, 2.0), box=6.0):
    """u0 = sum_k A_k exp(-|x - x_k|^2 / w_k^2); v0 = 0."""
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0x6A55)
    u = np.zeros((n, n))
    params = []
    for _ in range(k):
        A = ps.uniform(*amp)
        w = ps.uniform(*width)
        cx, cy = ps.uniform(-box, box), ps.uniform(-box, box)
        u += A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / w ** 2)
        params.append((A, w, cx, cy))
    return u, np.zeros((n, n)), params


def kink_ring_2d(n, L, seed, traj, radius=(2.0, 6.0), width=(0.5, 1.5)):
    """u0 = tanh((R - r)/w) centred at 0; v0 = 0."""
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0x4B12)
    R = ps.uniform(*radius)
    w = ps.uniform(*width)
    r = np.sqrt(X ** 2 + Y ** 2)
    return np.tanh((R - r) / w), np.zeros((n, n)), {"R": R, "w": w}


def nlse_packets_2d(n, L, seed, traj, kmin=2, kmax=4, amp=(0.5, 1.5), width=1.0, box=6.0):
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0x4E15)
    K = ps.integer(kmin, kmax)
    psi = np.zeros((n, n), np.complex128)
    params = []
    for _ in range(K):
        A = ps.uniform(*amp)
        phi = ps.uniform(0.0, 2 * np.pi)
        vx, vy = float(ps.normal()), float(ps.normal())
        cx, cy = ps.uniform(-box, box), ps.uniform(-box, box)
        env = A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / width ** 2)
        psi += env * np.exp(1j * (vx * (X - cx) + vy * (Y - cy) + phi))
        params.append((A, phi, vx, vy, cx, cy))
    return psi, params


def bubbles_3d(n, L, seed, traj, kmin=2, kmax=5, amp=(0.5, 2.0), width=(0.8, 2.0), box=4.0):
    x = np.linspace(-L, L, n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ps = ParamStream(seed, traj, tag=0x3DB0)
    K = ps.integer(kmin, kmax)
    u = np.zeros((n, n, n))
    for _ in range(K):
        A = ps.uniform(*amp)
        w = ps.uniform(*width)
        c = ps.uniform(-box, box, size=3)
        u += A * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / w ** 2)
    return u, np.zeros_like(u)


def constant_fields(shape, m=1.0, c=1.0):
    return np.full(shape, float(m)), np.full(shape, float(c))
