"""Writes tests/golden/sampler_forms.npz: outputs of the reference's initial-condition and
coefficient-field functions (scripts/real_sampler.py, scripts/m_fields_2d.py,
scripts/c_fields_2d.py) evaluated at the parameters this library's samplers draw, for the
field classes whose form is a function of explicit parameters:

  kink ring     RealWaveSampler.ring_soliton('klein_gordon', radius=R, width=w)
                (real_sampler.py:435-467)
  c classes     constant, periodic_structure, piecewise_constant, sign_changing_mass
                (c_fields_2d.py:12-52)
  m classes     constant_mass, piecewise_constant_mass, gradient_based_mass,
                phase_shifted_mass, topological_mass (m_fields_2d.py:14-82), the last three
                on a conductivity field stored with them

The samplers that draw inside the reference functions are checked by replay: this
library's sampler runs with a recording wrapper around its Philox stream, and the
reference function then runs with numpy's legacy np.random calls answered, in order, by
the recorded draws (the two must make the same calls in the same order), on the same
explicit parameters:

  m classes     defects, quasiperiodic, multiscale (generate_m_fields draws,
                m_fields_2d.py:196-226, and localized_mass_defects / quasi_periodic_mass /
                multiscale_mass, :83-146)
  c classes     layered, waveguide, quasiperiodic, turbulent (generate_c_fields draws,
                c_fields_2d.py:184-220, and the class functions, :53-168)
  kink_field    RealWaveSampler.kink_field with its anisotropic GRF velocity
                (real_sampler.py:223-271, 48-65)
  multi_soliton NLSEPhenomenonSampler.multi_soliton_state (nlse_sampler.py:102-219),
                every system type, arrangement and phase pattern of the launcher's space
  soliton       NLSEPhenomenonSampler.fundamental_soliton at explicit parameters
                (nlse_sampler.py:43-100), the four system types

Run from the repository root in a container that has the reference sources under
/root/reference/scripts:

    python tests/golden/make_sampler_golden.py
"""
import contextlib
import os
import sys

import numpy as np

sys.dont_write_bytecode = True  # leave no bytecode next to the reference sources

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "nlsolvers-datagen_amd"))
sys.path.insert(0, "/root/reference/scripts")

import c_fields_2d as rc  # noqa: E402  (reference)
import m_fields_2d as rm  # noqa: E402  (reference)
from nlse_sampler import NLSEPhenomenonSampler  # noqa: E402  (reference)
from real_sampler import RealWaveSampler  # noqa: E402  (reference)

from nls import samplers as S  # noqa: E402
from nls.fields import c_field, m_field  # noqa: E402
from nls.samplers import ParamStream, kink_ring_2d  # noqa: E402

N, L = 24, 10.0
SEEDS = [(3, 0), (3, 1), (11, 4)]
# replayed samplers: (seed, trajectory) pairs and grid
REPLAY_SEEDS = [(5, t) for t in range(6)]
NR, LR = 32, 6.0


class Recorder:
    """Forwards the sampler's draws to a ParamStream and records each returned value."""

    def __init__(self, ps):
        self.ps, self.log = ps, []

    def _rec(self, v):
        self.log.append(np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
        return v

    def uniform(self, lo=0.0, hi=1.0, size=None):
        return self._rec(self.ps.uniform(lo, hi, size=size))

    def integer(self, lo, hi):
        return self._rec(self.ps.integer(lo, hi))

    def normal(self, loc=0.0, scale=1.0, size=None):
        return self._rec(self.ps.normal(loc, scale, size=size))

    def rand(self, *shape):
        return self._rec(self.ps.rand(*shape))

    def randn(self, *shape):
        return self._rec(self.ps.randn(*shape))

    def randint(self, lo, hi):
        return self._rec(self.ps.randint(lo, hi))

    def random(self):
        return self._rec(self.ps.random())


@contextlib.contextmanager
def replay(values):
    """np.random's legacy calls answer with `values` in order (shapes checked)."""
    q = list(values)
    names = ("uniform", "randint", "randn", "rand", "normal", "random")
    saved = {k: getattr(np.random, k) for k in names}

    def pop(size=None):
        assert q, "the reference made more draws than the sampler"
        v = q.pop(0)
        if size is not None:
            shape = (size,) if np.ndim(size) == 0 else tuple(size)
            assert np.shape(v) == shape, (np.shape(v), shape)
        return v

    np.random.uniform = lambda low=0.0, high=1.0, size=None: pop(size)
    np.random.randint = lambda low, high=None, size=None: pop(size)
    np.random.randn = lambda *shape: pop(shape if shape else None)
    np.random.rand = lambda *shape: pop(shape if shape else None)
    np.random.normal = lambda loc=0.0, scale=1.0, size=None: pop(size)
    np.random.random = lambda size=None: pop(size)
    try:
        yield q
    finally:
        for k, f in saved.items():
            setattr(np.random, k, f)
    assert not q, f"{len(q)} recorded draws left unused"


def replayed_fields(out):
    """The field classes drawn inside the reference functions, by replay."""
    for s, t in SEEDS:
        tag = f"{s}_{t}"
        for kind in ("defects", "quasiperiodic", "multiscale"):
            rec = Recorder(ParamStream(s, t, tag=0x3A55))
            _, q = m_field(kind, N, L, s, t, rng=rec)
            with replay(rec.log):
                ref, _ = rm.generate_m_fields(N, L, num_fields=1, field_types=[kind])
            out[f"mr_{kind}_{tag}"] = ref[0]
        for kind in ("layered", "waveguide", "quasiperiodic", "turbulent"):
            rec = Recorder(ParamStream(s, t, tag=0xC0C0))
            c_field(kind, N, L, s, t, rng=rec)
            # generate_c_fields first picks the class with np.random.choice
            saved = np.random.choice
            np.random.choice = lambda a, *args, **kw: kind
            try:
                with replay(rec.log):
                    ref, _ = rc.generate_c_fields(N, L, num_fields=1, field_types=[kind])
            finally:
                np.random.choice = saved
            out[f"cr_{kind}_{tag}"] = ref[0]


def replayed_phenomena(out):
    """kink_field and multi_soliton_state by replay; fundamental_soliton at explicit
    parameters for every system type."""
    out["replay_n"], out["replay_L"] = np.array(NR), np.array(LR)
    out["replay_seeds"] = np.array(REPLAY_SEEDS)
    for s, t in REPLAY_SEEDS:
        tag = f"{s}_{t}"
        rec = Recorder(ParamStream(s, t, tag=0x4B1F))
        _, _, p = S.kink_field_2d(NR, LR, s, t, rng=rec)
        with replay(rec.log):
            u, v = RealWaveSampler(NR, NR, LR).kink_field(
                p["system_type"], winding_x=p["winding_x"], winding_y=p["winding_y"],
                width_range=p["width_range"], randomize_positions=p["randomize_positions"])
        out[f"kink_u_{tag}"], out[f"kink_v_{tag}"] = u, v
        rec = Recorder(ParamStream(s, t, tag=0x50A2))
        _, p = S.multi_soliton_2d(NR, LR, s, t, rng=rec, normalize=False)
        kw = {k: p[k] for k in ("amplitude_range", "width_range", "phase_pattern",
                                "arrangement", "coherence", "velocity_scale", "chirp_range",
                                "aspect_ratio_range")}
        with replay(rec.log):
            psi = NLSEPhenomenonSampler(NR, NR, LR).multi_soliton_state(p["system_type"], **kw)
        out[f"solitons_{tag}"] = psi
    # every system type, arrangement and phase pattern at least once
    for k, (st, arr, ph) in enumerate([("cubic", "circular", "alternating"),
                                       ("cubic_quintic", "linear", "synchronized"),
                                       ("saturable", "lattice", "random"),
                                       ("glasner_allen_flowers", "random", "vortex"),
                                       ("cubic", "random", "partial_coherence")]):
        rec = Recorder(ParamStream(9, k, tag=0x50A2))
        _, p = S.multi_soliton_2d(NR, LR, 9, k, system_type=st, arrangement=arr,
                                  phase_pattern=ph, velocity_scale=0.5, rng=rec,
                                  normalize=False)
        kw = {kk: p[kk] for kk in ("amplitude_range", "width_range", "phase_pattern",
                                   "arrangement", "coherence", "velocity_scale", "chirp_range",
                                   "aspect_ratio_range")}
        with replay(rec.log):
            psi = NLSEPhenomenonSampler(NR, NR, LR).multi_soliton_state(st, **kw)
        out[f"solitons_case{k}"] = psi
    X, Y = S.grid_2d(NR, LR)
    for k, st in enumerate(("cubic", "cubic_quintic", "saturable", "glasner_allen_flowers")):
        kw = dict(amplitude=1.3, width=0.9, position=(0.7, -1.1), phase=0.4,
                  velocity=(0.3, -0.6), Lambda=0.08, chirp_factor=0.05, aspect_ratio=1.3,
                  orientation=0.6, order=1 + (k % 2))
        out[f"soliton_{st}"] = NLSEPhenomenonSampler(NR, NR, LR).fundamental_soliton(st, **kw)


def main():
    out = {"n": np.array(N), "L": np.array(L), "seeds": np.array(SEEDS)}
    for s, t in SEEDS:
        tag = f"{s}_{t}"
        _, _, p = kink_ring_2d(N, L, s, t)
        u, v = RealWaveSampler(N, N, L).ring_soliton(
            system_type="klein_gordon", amplitude=1.0, radius=p["R"], width=p["w"],
            position=(0.0, 0.0), velocity=0.0, ring_type="expanding")
        out[f"ring_u_{tag}"], out[f"ring_v_{tag}"] = u, v
        # conductivity classes with explicit parameters
        _, q = c_field("periodic_structure", N, L, s, t)
        out[f"c_periodic_structure_{tag}"] = rc.periodic_structure(
            N, L, amplitude=q["amplitude"], frequency=q["frequency"])
        _, q = c_field("piecewise_constant", N, L, s, t)
        out[f"c_piecewise_constant_{tag}"] = rc.piecewise_constant(
            N, L, num_layers=q["num_layers"], contrast_factor=q["contrast_factor"])
        _, q = c_field("sign_changing_mass", N, L, s, t)
        out[f"c_sign_changing_mass_{tag}"] = rc.sign_changing_mass(
            N, L, regions=q["regions"], scale=q["scale"], sharpness=q["sharpness"])
        out[f"c_constant_{tag}"] = rc.constant(N, L)
        # cubic-coefficient classes; the c-derived ones on a sign-changing c (so the
        # topological class's sign(lap c - lambda c) changes across the grid)
        c, _ = c_field("sign_changing_mass", N, L, s, t)
        out[f"c_for_m_{tag}"] = c
        _, q = m_field("piecewise", N, L, s, t)
        out[f"m_piecewise_{tag}"] = rm.piecewise_constant_mass(
            N, L, m1=1.0, m2=q["m2"], boundary_type=q["boundary_type"],
            boundary_param=q["boundary_param"], smooth_width=q["smooth_width"])
        _, q = m_field("gradient", N, L, s, t, c=c)
        out[f"m_gradient_{tag}"] = rm.gradient_based_mass(N, L, c, gamma=q["gamma"],
                                                           epsilon=q["epsilon"])
        _, q = m_field("phase", N, L, s, t, c=c)
        out[f"m_phase_{tag}"] = rm.phase_shifted_mass(N, L, c, delta=q["delta"],
                                                       shift_fraction=q["shift_fraction"])
        _, q = m_field("topological", N, L, s, t, c=c)
        out[f"m_topological_{tag}"] = rm.topological_mass(N, L, c, eta=q["eta"],
                                                          lambda_param=q["lambda_param"])
        out[f"m_constant_{tag}"] = rm.constant_mass(N, L)
    replayed_fields(out)
    replayed_phenomena(out)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampler_forms.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
