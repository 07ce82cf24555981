"""Coefficient fields m(x) (cubic term) and c(x) (conductivity) on the n x n grid.

The classes and their functional forms are those of the reference generators
(scripts/m_fields_2d.py:8-146 with the class dispatch at :147-238, and
scripts/c_fields_2d.py:8-168 with the dispatch at :169-248); the names are the ones
the dataset configurations list under m_classes / c_classes.  Random parameters
are drawn from the Philox ParamStream keyed by (seed, trajectory), so a
trajectory's fields do not depend on how many others were drawn (the reference uses
the global legacy np.random stream, whose draws are therefore not reproduced).
Every generator returns (field, params).  Host-side numpy/scipy code: fields are
built once per trajectory, outside the time stepping.
"""
import numpy as np
from scipy.ndimage import gaussian_filter, laplace
from scipy.spatial.distance import cdist

from .samplers import ParamStream, grid_2d

M_CLASSES = ("constant", "piecewise", "gradient", "phase", "topological", "defects",
             "quasiperiodic", "multiscale")
C_CLASSES = ("constant", "periodic_structure", "piecewise_constant", "sign_changing_mass",
             "layered", "waveguide", "quasiperiodic", "turbulent")
_GOLDEN = (1.0 + np.sqrt(5.0)) / 2.0


def _rescale(p, base):
    lo, hi = p.min(), p.max()
    return base * (p - lo) / (hi - lo) if hi > lo else np.full_like(p, base)


def _quasi_waves(X, Y, ps, num_waves, amp_lo, amp_hi, base, scale_amp):
    p = np.full(X.shape, float(base))
    for i in range(num_waves):
        ang = i * np.pi / num_waves
        s = _GOLDEN ** i
        kx, ky = np.cos(ang) * s, np.sin(ang) * s
        a = ps.uniform(amp_lo, amp_hi) * (base if scale_amp else 1.0)
        ph = ps.uniform(0.0, 2 * np.pi)
        p += a * np.cos(kx * X + ky * Y + ph)
    return p


# ---------------------------------------------------------------- c(x)
def c_field(kind, n, L, seed, traj, base=1.0, rng=None):
    """One conductivity field of class `kind` (c_fields_2d.py:169-248 parameter ranges).
    rng (default: the trajectory's Philox stream) supplies the draws, in the order the
    reference's generator makes its np.random calls for the classes drawn from loops."""
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0xC0C0) if rng is None else rng
    if kind == "constant":
        return np.ones((n, n)), {"type": "constant"}
    if kind == "periodic_structure":
        amp, freq = ps.uniform(0.2, 0.5), ps.integer(1, 2)
        k = np.pi * freq / L
        return base * (1 + amp * np.sin(k * X) * np.sin(k * Y)), \
            {"type": kind, "amplitude": amp, "frequency": freq}
    if kind == "piecewise_constant":
        layers, contrast = ps.integer(2, 4), ps.uniform(1.5, 2.5)
        p = np.full((n, n), float(base))
        w = 2 * L / layers
        for i in range(1, layers, 2):
            p[(X >= -L + i * w) & (X < -L + (i + 1) * w)] = base * contrast
        return p, {"type": kind, "num_layers": layers, "contrast_factor": contrast}
    if kind == "sign_changing_mass":
        regions = "checkerboard" if ps.uniform() < 0.5 else "half_space"
        scale, sharp = 2, ps.uniform(3.0, 6.0)
        if regions == "checkerboard":
            cell = L / scale
            p = base * np.tanh(sharp * np.sin(np.pi * X / cell) * np.sin(np.pi * Y / cell))
        else:
            p = base * np.tanh(sharp * X / L)
        return p, {"type": kind, "regions": regions, "scale": scale, "sharpness": sharp}
    if kind == "layered":
        layers = ps.integer(2, 5)
        a_lo, a_hi = ps.uniform(0.1, 0.3), ps.uniform(0.4, 0.8)
        f_lo, f_hi = ps.uniform(1.0, 3.0), ps.uniform(5.0, 15.0)
        p = np.full((n, n), float(base))
        for _ in range(layers):
            th = ps.uniform(0.0, np.pi)
            d = np.cos(th) * X + np.sin(th) * Y
            p += ps.uniform(a_lo, a_hi) * np.sin(ps.uniform(f_lo, f_hi) * d
                                                  + ps.uniform(0.0, 2 * np.pi))
        return _rescale(p, base), {"type": kind, "num_layers": layers}
    if kind == "waveguide":
        w_lo, w_hi, gamp = ps.uniform(0.1, 0.3), ps.uniform(0.4, 0.8), ps.uniform(0.5, 2.0)
        guides = ps.integer(3, 11)
        p = np.full((n, n), float(base))
        pts = np.column_stack([X.ravel(), Y.ravel()])
        for _ in range(guides):
            line = ps.uniform() < 0.5
            width = ps.uniform(w_lo, w_hi)
            if line:  # straight guide through a random point
                x0, y0, th = ps.uniform(-L, L), ps.uniform(-L, L), ps.uniform(0, 2 * np.pi)
                t = np.linspace(-1.5 * L, 1.5 * L, 100)
                cx, cy = x0 + t * np.cos(th), y0 + t * np.sin(th)
            else:  # ellipse
                a, b, ph = ps.uniform(0.5, 2.0), ps.uniform(0.5, 2.0), ps.uniform(0, 2 * np.pi)
                t = np.linspace(0, 2 * np.pi, 100)
                cx, cy = a * np.cos(t + ph), b * np.sin(t)
            dist = cdist(pts, np.column_stack([cx, cy])).min(axis=1).reshape(n, n)
            p = np.maximum(p, gamp * np.exp(-dist ** 2 / (2 * width ** 2)))
        return p, {"type": kind, "num_guides": guides, "guide_amp": gamp}
    if kind == "quasiperiodic":
        waves = ps.integer(3, 7)
        p = _quasi_waves(X, Y, ps, waves, ps.uniform(0.1, 0.3), ps.uniform(0.4, 0.8), base, False)
        return _rescale(p, base), {"type": kind, "num_waves": waves}
    if kind == "turbulent":
        inten, s_lo, s_hi = ps.uniform(0.3, 0.8), ps.uniform(1.0, 3.0), ps.uniform(10.0, 30.0)
        beta, octaves = ps.uniform(1.0, 3.0), ps.integer(3, 7)
        f = np.zeros((n, n))
        for o in range(octaves):
            sc = s_hi / 2 ** o
            if sc < s_lo:
                break
            f += sc ** beta * gaussian_filter(ps.normal(size=(n, n)), sc)
        f -= f.min()
        f /= f.max()
        return base * np.exp(inten * (f - 0.5)), {"type": kind, "intensity": inten,
                                                  "beta": beta, "num_octaves": octaves}
    raise ValueError(f"unknown c class {kind!r}; known: {C_CLASSES}")


# ---------------------------------------------------------------- m(x)
def m_field(kind, n, L, seed, traj, c=None, m0=1.0, rng=None):
    """One cubic-coefficient field of class `kind` (m_fields_2d.py:147-238 parameter
    ranges); gradient / phase / topological derive from the conductivity c.  rng as in
    c_field."""
    X, Y = grid_2d(n, L)
    ps = ParamStream(seed, traj, tag=0x3A55) if rng is None else rng
    if kind in ("gradient", "phase", "topological") and c is None:
        kind = "constant"  # as the reference does without a c field
    if kind == "constant":
        return np.full((n, n), float(m0)), {"type": "constant", "m0": m0}
    if kind == "piecewise":
        m2 = ps.uniform(1.5, 3.0) * m0
        bt = ("circle", "square", "horizontal", "vertical", "diagonal")[ps.integer(0, 4)]
        bp, sw = ps.uniform(0.3, 0.7), ps.uniform(0.01, 0.1)
        b = {"circle": np.sqrt(X ** 2 + Y ** 2) - bp * L,
             "square": np.maximum(np.abs(X), np.abs(Y)) - bp * L,
             "horizontal": Y, "vertical": X, "diagonal": X + Y}[bt]
        return m0 + (m2 - m0) * 0.5 * (1 + np.tanh(b / (sw * L))), \
            {"type": kind, "m2": m2, "boundary_type": bt, "boundary_param": bp, "smooth_width": sw}
    if kind == "gradient":
        gam, eps = ps.uniform(0.5, 2.0), ps.uniform(1e-3, 1e-1)
        gx, gy = np.zeros_like(c), np.zeros_like(c)
        gx[1:-1, :] = (c[2:, :] - c[:-2, :]) / 2
        gy[:, 1:-1] = (c[:, 2:] - c[:, :-2]) / 2
        g2 = gx ** 2 + gy ** 2
        return m0 * (1 + gam * g2 / (g2 + eps ** 2)), {"type": kind, "gamma": gam, "epsilon": eps}
    if kind == "phase":
        delta, frac = ps.uniform(0.3, 1.0), ps.uniform(0.02, 0.1)
        k = max(1, int(frac * n))
        dx = np.roll(c, k, axis=0) - np.roll(c, -k, axis=0)
        dy = np.roll(c, k, axis=1) - np.roll(c, -k, axis=1)
        mag = np.sqrt(dx ** 2 + dy ** 2)
        mx = np.max(np.abs(mag))
        return m0 * (1 + delta * (mag / mx if mx > 0 else mag)), \
            {"type": kind, "delta": delta, "shift_fraction": frac}
    if kind == "topological":
        eta, lam = ps.uniform(0.5, 1.0), ps.uniform(0.3, 0.7)
        t = gaussian_filter(np.sign(laplace(c) - lam * c), sigma=1.0)
        mx = np.max(np.abs(t))
        return m0 * (1 + eta * (t / mx if mx > 0 else t)), \
            {"type": kind, "eta": eta, "lambda_param": lam}
    if kind == "defects":
        num = ps.integer(5, 19)
        s_lo, s_hi = ps.uniform(-0.5, -0.1), ps.uniform(0.5, 1.0)
        w_lo, w_hi = ps.uniform(0.03, 0.08), ps.uniform(0.1, 0.3)
        m = np.full((n, n), float(m0))
        for _ in range(num):
            x0, y0 = ps.uniform(-L, L), ps.uniform(-L, L)
            st, w = ps.uniform(s_lo, s_hi) * m0, ps.uniform(w_lo, w_hi) * L
            m += st * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2 * w ** 2))
        return np.maximum(m, 0.1 * m0), {"type": kind, "num_defects": num}
    if kind == "quasiperiodic":
        waves = ps.integer(3, 7)
        m = _quasi_waves(X, Y, ps, waves, ps.uniform(0.1, 0.3), ps.uniform(0.4, 0.8), m0, True)
        return np.maximum(m, 0.1 * m0), {"type": kind, "num_waves": waves}
    if kind == "multiscale":
        scales = ps.integer(3, 5)
        s_lo, s_hi = ps.uniform(1.0, 3.0), ps.uniform(8.0, 20.0)
        a_lo, a_hi = ps.uniform(0.1, 0.3), ps.uniform(0.4, 0.8)
        m = np.full((n, n), float(m0))
        for sc in np.logspace(np.log10(s_lo), np.log10(s_hi), scales):
            sm = gaussian_filter(ps.normal(size=(n, n)), sigma=sc)
            m += ps.uniform(a_lo, a_hi) * m0 * sm / np.max(np.abs(sm))
        return np.maximum(m, 0.1 * m0), {"type": kind, "num_scales": scales}
    raise ValueError(f"unknown m class {kind!r}; known: {M_CLASSES}")


# ---------------------------------------------------------------- the same classes on the GPU
def _field_device(family, kind, n, L, seed, traj, base, c=None):
    import ctypes as C
    from ._lib import call
    from .device import DeviceArray
    out = DeviceArray((n, n), np.float64)
    call("nls_field_sample", family.encode(), kind.encode(), int(n), float(L), int(seed),
         int(traj), float(base), C.c_void_p(c.ptr) if c is not None else None,
         C.c_void_p(out.ptr), None)
    return out


def c_field_device(kind, n, L, seed, traj, base=1.0):
    """c_field evaluated on the GPU (nls_field_sample): the same draws from the
    trajectory's Philox stream, the field left in device memory (a DeviceArray)."""
    return _field_device("c", kind, n, L, seed, traj, base)


def m_field_device(kind, n, L, seed, traj, c=None, m0=1.0):
    """m_field on the GPU; c is the device conductivity field (DeviceArray) the gradient,
    phase and topological classes derive from (constant without it)."""
    return _field_device("m", kind, n, L, seed, traj, m0, c)

