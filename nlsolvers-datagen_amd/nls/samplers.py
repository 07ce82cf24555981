"""Seeded initial-condition samplers.

The benchmark configurations' samplers (shapes from the reference, parameter laws as
the configurations state them, centres and radii scaled with the half-width L so that
the structure lies inside any domain; L = 10 (2D) and L = 8 (3D) give the stated laws):
  gaussian_sum_2d   sum of Gaussians A exp(-|x-x0|^2/w^2)  (util.hpp:94-125 profile),
                    centres U(-0.6 L, 0.6 L)^2
  kink_ring_2d      tanh((R - r)/w), R ~ U(0.2 L, 0.6 L)     (real_sampler.py:464-467)
  nlse_packets_2d   Gaussian packets A exp(-|x-x0|^2/w^2) exp(i(v.(x-x0) + phi))
                                                           (nlse_sampler.py:43-100 form)
  bubbles_3d        sum of 3D Gaussian bubbles A exp(-|x-x0|^2/w^2)  (util.hpp:94-125
                    profile; the configurations' 3D Gaussian bubbles), centres
                    U(-0.5 L, 0.5 L)^3
The reference's own phenomenon samplers, with its draw sequence and parameter spaces:
  kink_field_2d     RealWaveSampler.kink_field (real_sampler.py:223-271) with the
                    anisotropic Gaussian-random-field velocity (:48-65)
  fundamental_soliton  NLSEPhenomenonSampler.fundamental_soliton (nlse_sampler.py:43-100)
  multi_soliton_2d  NLSEPhenomenonSampler.multi_soliton_state (nlse_sampler.py:102-219)
Grids are linspace(-L, L, n) with indexing="ij" (real_sampler.py:17-25).

Every random parameter is a Philox4x32-10 draw keyed by the seed, with counter
(slot, trajectory, tag, 0), so a trajectory's initial data do not depend on how
many trajectories are drawn or in which order (the reference uses the global
legacy np.random stream, real_launcher_2d.py:390-395).  ParamStream offers the
legacy-numpy calls the reference samplers make (rand, randn, randint, uniform,
normal, random), and the reference-form samplers make those calls in the
reference's order.
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

    def normal(self, loc=0.0, scale=1.0, size=None):
        """Normal draws by Box-Muller; loc + scale z (z itself for loc 0, scale 1)."""
        u1 = self.uniform(size=size)
        u2 = self.uniform(size=size)
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)
        return z if (loc == 0.0 and scale == 1.0) else loc + scale * z

    # the legacy np.random calls of the reference samplers
    def rand(self, *shape):
        return self.uniform(size=shape) if shape else self.uniform()

    def randn(self, *shape):
        return self.normal(size=shape) if shape else self.normal()

    def randint(self, lo, hi):
        """Integer in [lo, hi) (np.random.randint)."""
        return self.integer(lo, hi - 1)

    def random(self):
        return self.uniform()


def grid_2d(n, L):
    x = np.linspace(-L, L, n)
    return np.meshgrid(x, x, indexing="ij")


def gaussian_sum_2d(n, L, seed, traj, k=3, amp=(0.5, 2.0), width=(0.5, 2.0), box=None):
    """u0 = sum_k A_k exp(-|x - x_k|^2 / w_k^2), centres U(-box, box)^2 with box = 0.6 L
    by default; v0 = 0."""
    box = 6.0 * L / 10.0 if box is None else box
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


def kink_ring_2d(n, L, seed, traj, radius=None, width=(0.5, 1.5)):
    """u0 = tanh((R - r)/w) centred at 0, R ~ U(0.2 L, 0.6 L) by default; v0 = 0 (the form
    of ring_soliton('klein_gordon', amplitude=1, velocity=0), real_sampler.py:435-467)."""
    radius = (2.0 * L / 10.0, 6.0 * L / 10.0) if radius is None else radius
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0x4B12)
    R = ps.uniform(*radius)
    w = ps.uniform(*width)
    r = np.sqrt(X ** 2 + Y ** 2)
    return np.tanh((R - r) / w), np.zeros((n, n)), {"R": R, "w": w}


def nlse_packets_2d(n, L, seed, traj, kmin=2, kmax=4, amp=(0.5, 1.5), width=1.0, box=None):
    """K ~ U{2..4} Gaussian packets, centres U(-box, box)^2 with box = 0.6 L by default."""
    box = 6.0 * L / 10.0 if box is None else box
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


def bubbles_3d(n, L, seed, traj, kmin=2, kmax=5, amp=(0.5, 2.0), width=(0.8, 2.0), box=None):
    """K ~ U{2..5} bubbles, centres U(-box, box)^3 with box = 0.5 L by default."""
    box = 4.0 * L / 8.0 if box is None else box
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


# ---------------------------------------------------------------- reference phenomena
# Parameter spaces the reference launchers draw from (valid_spaces_real.py:33-41,
# valid_spaces_complex.py:11-21); every listed value is equally likely.
KINK_FIELD_SPACE = {
    "system_type": ["sine_gordon", "double_sine_gordon", "hyperbolic_sine_gordon", "phi4",
                    "klein_gordon"],
    "winding_x": list(range(-4, 5)),
    "winding_y": list(range(-4, 5)),
    "width_range": [(a, b) for a in (0.3, 0.5, 0.7) for b in (1.5, 2.0, 3.0)],
    "randomize_positions": [True, False],
}
MULTI_SOLITON_SPACE = {
    "system_type": ["cubic", "cubic_quintic", "saturable", "glasner_allen_flowers"],
    "width_range": [(0.5, 1.0), (1.0, 1.5), (1.5, 2.0)],
    "amplitude_range": [(0.5, 1.0), (1.0, 1.5), (1.5, 2.0)],
    "phase_pattern": ["random", "alternating", "synchronized", "vortex"],
    "arrangement": ["linear", "circular", "random", "lattice"],
    "coherence": [0.2, 0.5, 0.8],
    "velocity_scale": [0.0, 0.5, 1.0],
    "chirp_range": [(-0.5, 0.0), (0.0, 0.5)],
    "aspect_ratio_range": [(1.0, 1.0), (1.0, 1.5)],
}


def space_params(space, seed, traj, tag):
    """One choice per key of a parameter space (the launchers' sample_phenomenon_params,
    real_launcher_2d.py:92-111), drawn from the trajectory's Philox stream."""
    ps = ParamStream(seed, traj, tag=tag)
    return {k: v[ps.integer(0, len(v) - 1)] for k, v in space.items()}


def anisotropic_grf(n, L, rng, length_scale=1.0, anisotropy_ratio=2.0, theta=30.0, power=2.0,
                    amplitude=1.0):
    """Gaussian random field with the rotated anisotropic spectrum
    exp(-((kx'/lx)^2 + (ky'/ly)^2)^(p/2)) applied to complex white noise by FFT, scaled to
    standard deviation `amplitude` (RealWaveSampler.anisotropic_grf, real_sampler.py:48-65,
    wavenumbers 2 pi fftfreq(n, 2L/n) of :27-29)."""
    k = 2 * np.pi * np.fft.fftfreq(n, d=2 * L / n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    theta_rad = np.deg2rad(theta)
    ell_x = length_scale * np.sqrt(anisotropy_ratio)
    ell_y = length_scale / np.sqrt(anisotropy_ratio)
    KX_rot = KX * np.cos(theta_rad) - KY * np.sin(theta_rad)
    KY_rot = KX * np.sin(theta_rad) + KY * np.cos(theta_rad)
    spectrum = np.exp(-((KX_rot / ell_x) ** 2 + (KY_rot / ell_y) ** 2) ** (power / 2))
    noise = rng.randn(n, n) + 1j * rng.randn(n, n)
    field = np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spectrum)).real
    return field / np.std(field) * amplitude


def kink_field_2d(n, L, seed, traj, winding_x=None, winding_y=None, width_range=None,
                  randomize_positions=None, rng=None):
    """Sine-Gordon kinks 4 atan(exp((x - x0)/w)) along x (|winding_x| of them, sign of
    winding_x) and along y, widths width_x (0.8 + 0.4 U) with width_x ~ U(width_range),
    positions U(-L, L) or evenly spaced, and a velocity v0 from the anisotropic GRF with
    length scale 2 mean(width_range) and standard deviation 0.1 max|u0|
    (RealWaveSampler.kink_field, real_sampler.py:223-271).  Parameters left None are drawn
    from the reference's kink_field space; rng (default: the trajectory's Philox stream)
    supplies the sampler's own draws in the reference's order."""
    pr = space_params(KINK_FIELD_SPACE, seed, traj, 0x4B1E)
    winding_x = pr["winding_x"] if winding_x is None else winding_x
    winding_y = pr["winding_y"] if winding_y is None else winding_y
    width_range = pr["width_range"] if width_range is None else tuple(width_range)
    randomize_positions = (pr["randomize_positions"] if randomize_positions is None
                           else randomize_positions)
    rng = ParamStream(seed, traj, tag=0x4B1F) if rng is None else rng
    X, Y = grid_2d(n, L)
    u0 = np.zeros_like(X)
    for winding, G in ((winding_x, X), (winding_y, Y)):
        if winding == 0:
            continue
        width = width_range[0] + (width_range[1] - width_range[0]) * rng.rand()
        if randomize_positions:
            positions = [L * (2 * rng.rand() - 1) for _ in range(abs(winding))]
        else:
            positions = [L * (-0.8 + 1.6 * i / (abs(winding))) for i in range(abs(winding))]
        sign = 1 if winding > 0 else -1
        for x0 in positions:
            kink_width = width * (0.8 + 0.4 * rng.rand())
            u0 += sign * 4 * np.arctan(np.exp((G - x0) / kink_width))
    v0 = anisotropic_grf(n, L, rng, length_scale=np.mean(width_range) * 2.0,
                         amplitude=np.max(np.abs(u0)) * 0.1)
    params = {"system_type": pr["system_type"], "winding_x": winding_x, "winding_y": winding_y,
              "width_range": width_range, "randomize_positions": randomize_positions}
    return u0, v0, params


def _sech(x):
    return 1.0 / np.cosh(x)


def fundamental_soliton(X, Y, L, system_type, amplitude=1.0, width=1.0, position=(0, 0),
                        phase=0.2, velocity=(0.0, 0.0), sigma1=1.0, sigma2=-0.1, kappa=1.0,
                        apply_envelope=True, envelope_width=0.7, Lambda=0.1, chirp_factor=0.0,
                        aspect_ratio=1.0, orientation=0.0, order=1):
    """A localized soliton profile (sech for the cubic NLSE, and the cubic-quintic,
    saturable and Glasner-Allen-Flowers profiles) with momentum, chirp and constant
    phase, on an elliptical rotated frame, optionally under the Gaussian envelope
    exp(-r^2 / (2 (0.7 L)^2)) (NLSEPhenomenonSampler.fundamental_soliton,
    nlse_sampler.py:29-100)."""
    x0, y0 = position
    vx, vy = velocity
    X_rot = (X - x0) * np.cos(orientation) + (Y - y0) * np.sin(orientation)
    Y_rot = -(X - x0) * np.sin(orientation) + (Y - y0) * np.cos(orientation)
    r_local = np.sqrt((X_rot / aspect_ratio) ** 2 + Y_rot ** 2)
    momentum_phase = vx * (X - x0) + vy * (Y - y0)
    chirp_phase = chirp_factor * r_local ** 2
    total_phase = momentum_phase + phase + chirp_phase
    if system_type == "cubic":
        if order == 1:
            profile = amplitude * _sech(r_local / width)
        else:
            profile = amplitude * _sech(r_local / width) ** order
    elif system_type == "cubic_quintic":
        beta = -sigma2 * amplitude ** 2 / sigma1
        if beta > 0:
            if order == 1:
                profile = amplitude * _sech(r_local / width) / np.sqrt(
                    1 + beta * _sech(r_local / width) ** 2)
            else:
                profile = amplitude * _sech(r_local / width) ** order / np.sqrt(
                    1 + beta * _sech(r_local / width) ** (2 * order))
        else:
            if order == 1:
                profile = amplitude * _sech(r_local / width)
            else:
                profile = amplitude * _sech(r_local / width) ** order
    elif system_type == "saturable":
        if order == 1:
            sech_term = _sech(r_local / width)
        else:
            sech_term = _sech(r_local / width) ** order
        denom = np.sqrt(1 + kappa * amplitude ** 2 * sech_term ** 2)
        profile = amplitude * sech_term / denom
    elif system_type == "glasner_allen_flowers":
        if order == 1:
            sech_term = _sech(np.sqrt(Lambda) * r_local)
            profile = amplitude * sech_term / np.sqrt(9 - 48 * Lambda * sech_term ** 2 + 31)
        else:
            sech_term = _sech(np.sqrt(Lambda) * r_local) ** order
            profile = amplitude * sech_term / np.sqrt(
                9 - 48 * Lambda * sech_term ** (2 / order) + 31)
    else:
        raise ValueError(f"Unknown system type: {system_type}")
    u = profile * np.exp(1j * total_phase)
    if apply_envelope:
        r = np.sqrt(X ** 2 + Y ** 2)
        ew = envelope_width * L
        u = u * np.exp(-r ** 2 / (2 * ew ** 2))
    return u


def multi_soliton_2d(n, L, seed, traj, system_type=None, amplitude_range=None, width_range=None,
                     position_variance=1.0, velocity_scale=None, phase_pattern=None,
                     arrangement=None, separation=5.0, sigma1=1.0, sigma2=-0.1, kappa=1.0,
                     apply_envelope=False, envelope_width=0.7, Lambda_range=(0.04, 0.14),
                     coherence=None, interaction_strength=0.5, cluster_levels=1,
                     order_range=(1, 2), chirp_range=None, aspect_ratio_range=None, rng=None,
                     normalize=True):
    """3..11 solitons (fundamental_soliton, apply_envelope False) placed linearly, on a
    circle, at random (N(0, L/4)) or on a lattice, with random / alternating /
    synchronized / vortex / partially coherent phases, each with its own amplitude,
    width, chirp, aspect ratio, orientation and velocity; components after the first are
    weighted by interaction_strength (NLSEPhenomenonSampler.multi_soliton_state,
    nlse_sampler.py:102-219).  Parameters left None are drawn from the reference's
    multi_soliton space; normalize divides by max|psi| as the complex launcher does
    (complex_launcher_2d.py:79-97)."""
    pr = space_params(MULTI_SOLITON_SPACE, seed, traj, 0x50A1)
    pick = lambda v, k: pr[k] if v is None else v  # noqa: E731
    system_type = pick(system_type, "system_type")
    amplitude_range = pick(amplitude_range, "amplitude_range")
    width_range = pick(width_range, "width_range")
    phase_pattern = pick(phase_pattern, "phase_pattern")
    arrangement = pick(arrangement, "arrangement")
    coherence = pick(coherence, "coherence")
    velocity_scale = pick(velocity_scale, "velocity_scale")
    chirp_range = pick(chirp_range, "chirp_range")
    aspect_ratio_range = pick(aspect_ratio_range, "aspect_ratio_range")
    rng = ParamStream(seed, traj, tag=0x50A2) if rng is None else rng
    X, Y = grid_2d(n, L)
    u = np.zeros_like(X, dtype=complex)
    n_solitons = rng.randint(3, 12)
    if arrangement == "linear":
        positions = [((i - (n_solitons - 1) / 2) * separation, 0) for i in range(n_solitons)]
    elif arrangement == "circular":
        positions = []
        for i in range(n_solitons):
            angle = 2 * np.pi * i / n_solitons
            positions.append((separation * np.cos(angle), separation * np.sin(angle)))
    elif arrangement == "random":
        positions = []
        for _ in range(n_solitons):
            x = rng.normal(0, position_variance * L / 4)
            y = rng.normal(0, position_variance * L / 4)
            positions.append((x, y))
    elif arrangement == "lattice":
        side = int(np.ceil(np.sqrt(n_solitons)))
        positions = []
        for i in range(side):
            for j in range(side):
                if len(positions) < n_solitons:
                    positions.append(((i - (side - 1) / 2) * separation,
                                      (j - (side - 1) / 2) * separation))
    else:
        raise ValueError(f"arrangement {arrangement!r} not built "
                         "(linear, circular, random, lattice)")
    if phase_pattern == "random":
        phases = rng.uniform(0, 2 * np.pi, size=n_solitons)
    elif phase_pattern == "alternating":
        phases = [i * np.pi for i in range(n_solitons)]
    elif phase_pattern == "synchronized":
        phases = np.zeros(n_solitons)
    elif phase_pattern == "vortex":
        center_x = np.mean([p[0] for p in positions])
        center_y = np.mean([p[1] for p in positions])
        phases = [np.arctan2(p[1] - center_y, p[0] - center_x) for p in positions]
    elif phase_pattern == "partial_coherence":
        base_phase = rng.uniform(0, 2 * np.pi)
        phases = []
        for _ in range(n_solitons):
            if rng.random() < coherence:
                phases.append(base_phase)
            else:
                phases.append(rng.uniform(0, 2 * np.pi))
    else:
        raise ValueError(f"phase_pattern {phase_pattern!r} unknown")
    for i, ((x0, y0), phase) in enumerate(zip(positions, phases)):
        amplitude = rng.uniform(*amplitude_range)
        width = rng.uniform(*width_range)
        Lambda = rng.uniform(*Lambda_range)
        order = rng.randint(*order_range)
        chirp = rng.uniform(*chirp_range)
        aspect_ratio = rng.uniform(*aspect_ratio_range)
        orientation = rng.uniform(0, 2 * np.pi)
        if velocity_scale > 0:
            if arrangement == "circular":
                angle = 2 * np.pi * i / n_solitons
                vx = -velocity_scale * np.cos(angle)
                vy = -velocity_scale * np.sin(angle)
            else:
                vx = rng.normal(0, velocity_scale)
                vy = rng.normal(0, velocity_scale)
        else:
            vx, vy = 0, 0
        component = fundamental_soliton(
            X, Y, L, system_type, amplitude=amplitude, width=width, position=(x0, y0),
            phase=phase, velocity=(vx, vy), sigma1=sigma1, sigma2=sigma2, kappa=kappa,
            apply_envelope=False, Lambda=Lambda, order=order, chirp_factor=chirp,
            aspect_ratio=aspect_ratio, orientation=orientation)
        if interaction_strength < 1.0 and i > 0:
            u = u + component * interaction_strength
        else:
            u += component
    if apply_envelope:
        r = np.sqrt(X ** 2 + Y ** 2)
        ew = envelope_width * L
        u = u * np.exp(-r ** 2 / (2 * ew ** 2))
    if normalize:
        u = u / np.max(np.abs(u))
    params = {"system_type": system_type, "amplitude_range": amplitude_range,
              "width_range": width_range, "phase_pattern": phase_pattern,
              "arrangement": arrangement, "coherence": coherence,
              "velocity_scale": velocity_scale, "chirp_range": chirp_range,
              "aspect_ratio_range": aspect_ratio_range, "n_solitons": n_solitons}
    return u, params


# ---------------------------------------------------------------- the phenomena on the GPU
def _space_choice(space, choices):
    return {k: v[int(i)] for (k, v), i in zip(space.items(), choices)}


def _phenomenon_device(name, n, L, seed, traj, normalize, complex_out):
    import ctypes as C
    from ._lib import call
    from .device import DeviceArray
    u = DeviceArray((n, n), np.complex128 if complex_out else np.float64)
    v = None if complex_out else DeviceArray((n, n), np.float64)
    ch = (C.c_int * 16)()
    cnt = C.c_int()
    call("nls_phenomenon_sample", name.encode(), int(n), float(L), int(seed), int(traj),
         int(bool(normalize)), C.c_void_p(u.ptr), C.c_void_p(v.ptr) if v is not None else None,
         ch, 16, C.byref(cnt), None)
    return u, v, list(ch[:cnt.value])


def kink_field_device(n, L, seed, traj):
    """kink_field_2d on the GPU (nls_phenomenon_sample): (u0, v0) as DeviceArrays and
    the chosen space parameters, the same draws and forms as the host sampler."""
    u, v, ch = _phenomenon_device("kink_field", n, L, seed, traj, False, False)
    return u, v, _space_choice(KINK_FIELD_SPACE, ch)


def multi_soliton_device(n, L, seed, traj, normalize=True):
    """multi_soliton_2d on the GPU: psi0 as a complex128 DeviceArray and the parameters
    (the chosen space values and n_solitons)."""
    u, _, ch = _phenomenon_device("multi_soliton", n, L, seed, traj, normalize, True)
    params = _space_choice(MULTI_SOLITON_SPACE, ch[:len(MULTI_SOLITON_SPACE)])
    params["n_solitons"] = ch[len(MULTI_SOLITON_SPACE)]
    return u, params

