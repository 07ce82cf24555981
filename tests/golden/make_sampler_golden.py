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

The classes built from internal np.random loops (m: defects, quasiperiodic, multiscale;
c: layered, waveguide, quasiperiodic, turbulent) are not covered.  Run from the repository
root in a container that has the reference sources under /root/reference/scripts:

    python tests/golden/make_sampler_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "nlsolvers-datagen_amd"))
sys.path.insert(0, "/root/reference/scripts")

import c_fields_2d as rc  # noqa: E402  (reference)
import m_fields_2d as rm  # noqa: E402  (reference)
from real_sampler import RealWaveSampler  # noqa: E402  (reference)

from nls.fields import c_field, m_field  # noqa: E402
from nls.samplers import kink_ring_2d  # noqa: E402

N, L = 24, 10.0
SEEDS = [(3, 0), (3, 1), (11, 4)]


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
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampler_forms.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
