"""Dataset generation launcher: the reference launchers' run loop on the fused batch.

The reference launchers (scripts/real_launcher_2d.py:260-306, scripts/complex_launcher_2d.py,
the 3D variants) draw an initial condition and the coefficient fields for each run, call
one solver executable per trajectory through .npy files, downsample the trajectory and
write one HDF5 file per run (real_launcher_2d.py:201-245).  Here a group of runs is
advanced together by the fused batch (KGE, NLSE) or the serving loop (stochastic phi^4),
and finished groups are written by the asynchronous HDF5 writer while the next group
computes.  Configuration comes from the launchers' command-line flags or from a dataset
YAML file with the keys of the reference's config_*.yaml (pde_type, boundary_condition,
num_trajectories, num_snapshots, nt, domain{Lx, Ly[, Lz], T}, grid{nx, ny[, nz]},
m_classes, c_classes) plus optional build keys (integrator, krylov_dim, mu2, seed,
phenomenon, batch, dr_x, dr_y, dr_strategy, gamma, sigma).
"""
import argparse
import os
import time
import uuid
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import fields as F
from . import samplers as S

_KGE = ("klein_gordon", "klein-gordon", "kge")
_NLSE = ("nonlinear schrodinger", "nonlinear_schrodinger", "nlse", "cubic")
_SP4 = ("phi4", "stochastic_phi4", "sp4")


@dataclass
class LaunchConfig:
    pde_type: str = "klein_gordon"
    num_runs: int = 1
    num_snapshots: int = 10
    nt: int = 500
    T: float = 5.0
    nx: int = 128
    ny: int = 128
    nz: int = 1
    dim: int = 2
    Lx: float = 10.0
    Ly: float = 10.0
    Lz: float = 10.0
    m_classes: List[str] = field(default_factory=lambda: ["constant"])
    c_classes: List[str] = field(default_factory=lambda: ["constant"])
    phenomenon: str = ""          # "" = the equation's default (PHENOMENA)
    integrator: str = ""          # gautschi | sewi | sv (KGE), strang (NLSE), damped | reference (phi^4)
    krylov_dim: int = 0           # 0: autotuner
    lanczos_variant: str = "full"  # full: textbook Lanczos (alpha_{m-1} computed);
                                   # reference: T[m-1][m-1] = 0 as lanczos.hpp:208-267
    mu2: float = 0.0
    gamma: float = 0.1
    sigma: float = 0.1
    seed: int = 0
    batch: int = 64               # trajectories per group
    dr_x: int = 128
    dr_y: int = 128
    dr_strategy: str = "none"     # none | interpolation | FFT
    output_dir: str = "results"
    run_id: Optional[str] = None
    sampler: str = "device"       # device: fields and phenomena drawn on the GPU
                                  # (nls_field_sample / nls_phenomenon_sample); host: nls.fields


def load_config(path, **overrides):
    """LaunchConfig from a dataset YAML file (reference config_*.yaml keys)."""
    import yaml
    with open(path) as f:
        y = yaml.safe_load(f)
    c = LaunchConfig()
    c.pde_type = str(y.get("pde_type", c.pde_type))
    bc = str(y.get("boundary_condition", "neumann")).lower()
    if bc not in ("neumann", "noflux"):
        raise ValueError(f"boundary_condition {bc!r}: only no-flux (Neumann) is built")
    c.num_runs = int(y.get("num_trajectories", c.num_runs))
    c.num_snapshots = int(y.get("num_snapshots", c.num_snapshots))
    c.nt = int(y.get("nt", c.nt))
    dom = y.get("domain", {})
    if "Lx" in dom:
        c.Lx = float(max(abs(v) for v in dom["Lx"]))
    c.Ly = float(max(abs(v) for v in dom["Ly"])) if "Ly" in dom else c.Lx
    c.Lz = float(max(abs(v) for v in dom["Lz"])) if "Lz" in dom else c.Lx
    if "T" in dom:
        c.T = float(dom["T"][1] - dom["T"][0])
    g = y.get("grid", {})
    c.nx = int(g.get("nx", c.nx))
    c.ny = int(g.get("ny", c.nx))
    if "nz" in g:
        c.nz, c.dim = int(g["nz"]), 3
    c.m_classes = list(y.get("m_classes", c.m_classes))
    c.c_classes = list(y.get("c_classes", c.c_classes))
    for k in ("integrator", "krylov_dim", "mu2", "seed", "phenomenon", "batch", "dr_x", "dr_y",
              "dr_strategy", "gamma", "sigma", "lanczos_variant"):
        if k in y:
            setattr(c, k, type(getattr(c, k))(y[k]))
    for k, v in overrides.items():
        if v is not None:
            setattr(c, k, v)
    return c


# ---------------------------------------------------------------- downsampling
def downsample_interpolation(u, target_shape, Lx, Ly):
    """Bilinear interpolation of each snapshot from linspace(-L, L, n) onto
    linspace(-L, L, target) (scripts/downsampling.py:68-103, RegularGridInterpolator
    'linear' on a regular grid)."""
    nt, nx, ny = u.shape
    tx, ty = target_shape

    def weights(n, t, L):
        # linear interpolation is separable: target point k takes (1 - w) of node i and
        # w of node i + 1, as one row of a (t, n) matrix with two non-zeros
        x = np.linspace(-L, L, n)
        xn = np.linspace(-L, L, t)
        i = np.clip(np.searchsorted(x, xn, side="right") - 1, 0, n - 2)
        w = (xn - x[i]) / (x[i + 1] - x[i])
        W = np.zeros((t, n))
        W[np.arange(t), i] = 1 - w
        W[np.arange(t), i + 1] += w
        return W

    Wx, Wy = weights(nx, tx, Lx), weights(ny, ty, Ly)
    return np.matmul(np.matmul(Wx, u), Wy.T).astype(u.dtype)


def downsample_fft(u, target_shape):
    """Spectral crop with orthonormal FFTs (scripts/downsampling.py:10-33)."""
    cplx = np.iscomplexobj(u)
    ft = np.fft.fftshift(np.fft.fft2(u, norm="ortho", axes=(-2, -1)), axes=(-2, -1))
    sx = (u.shape[-2] - target_shape[0]) // 2
    sy = (u.shape[-1] - target_shape[1]) // 2
    ft = ft[..., sx:sx + target_shape[0], sy:sy + target_shape[1]]
    out = np.fft.ifft2(np.fft.ifftshift(ft, axes=(-2, -1)), norm="ortho", axes=(-2, -1))
    return (out if cplx else out.real).astype(u.dtype)


def downsample(u, cfg):
    if cfg.dr_strategy == "none" or cfg.dim == 3:
        return u
    if cfg.dr_strategy == "interpolation":
        return downsample_interpolation(u, (cfg.dr_x, cfg.dr_y), cfg.Lx, cfg.Ly)
    if cfg.dr_strategy == "FFT":
        return downsample_fft(u, (cfg.dr_x, cfg.dr_y))
    raise ValueError(f"unknown downsampling strategy {cfg.dr_strategy!r}")


# ---------------------------------------------------------------- initial data
# The initial data each equation can be given: the reference launchers' phenomena that
# are built (kink_field: real_launcher_2d.py:318 default, RealWaveSampler.kink_field;
# multi_soliton: complex_launcher_2d.py:280 default, multi_soliton_state) and the
# benchmark configurations' samplers.  The first name is the default; any other name is
# refused (nothing is substituted under a name that was asked for).
PHENOMENA = {
    "real2d": ("kink_field", "kink_ring", "gaussians"),
    "nlse2d": ("multi_soliton", "packets"),
    "real3d": ("bubbles",),
}
_ALIASES = {"gaussian": "gaussians", "gaussian_sum": "gaussians", "kink_ring_2d": "kink_ring",
            "multi_soliton_state": "multi_soliton", "gaussian_packets": "packets",
            "bubbles_3d": "bubbles", "gaussian_bubbles": "bubbles"}


def resolve_phenomenon(name, family):
    """Canonical phenomenon name for an equation family ("real2d", "nlse2d", "real3d")."""
    built = PHENOMENA[family]
    if not name:
        return built[0]
    key = _ALIASES.get(name.lower(), name.lower())
    if key not in built:
        raise ValueError(f"phenomenon {name!r} is not built for this equation "
                         f"(built: {', '.join(built)})")
    return key


# ---------------------------------------------------------------- the launcher
class BatchLauncher:
    def __init__(self, cfg: LaunchConfig):
        self.cfg = cfg
        if cfg.sampler not in ("device", "host"):
            raise ValueError(f"sampler {cfg.sampler!r}: device or host")
        self.run_id = cfg.run_id or str(uuid.uuid4())[:8]
        self.h5_dir = os.path.join(cfg.output_dir, "hdf5")
        os.makedirs(self.h5_dir, exist_ok=True)
        pt = cfg.pde_type.lower()
        if pt in _KGE:
            self.kind = "kge"
        elif pt in _NLSE:
            self.kind = "nlse"
        elif pt in _SP4:
            self.kind = "sp4"
        else:
            raise ValueError(f"pde_type {cfg.pde_type!r} has no integrator here "
                             "(Klein-Gordon, Nonlinear Schrodinger, phi4)")
        if cfg.nx != cfg.ny or (cfg.dim == 3 and cfg.nz != cfg.nx):
            raise ValueError("grids must be square / cubic (as the reference solvers assume)")
        if cfg.dim == 3 and self.kind == "nlse":
            raise ValueError("3D NLSE datasets are not built here (the fused NLSE batch runs "
                             "them; no 3D NLSE sampler is)")
        family = "real3d" if cfg.dim == 3 else ("nlse2d" if self.kind == "nlse" else "real2d")
        self.phenomenon = resolve_phenomenon(cfg.phenomenon, family)
        if cfg.dim == 3 and (set(cfg.m_classes) | set(cfg.c_classes)) - {"constant"}:
            raise ValueError("3D coefficient-field classes other than 'constant' are not "
                             "built (the reference's field generators are 2D)")
        self.files = []

    def path(self, idx):
        return os.path.join(self.h5_dir, f"run_{self.run_id}_{idx:04d}.h5")

    def sample(self, idx):
        """(u0, v0, m, c, phenomenon params, m class, c class) of run idx."""
        cfg, n, ph = self.cfg, self.cfg.nx, self.phenomenon
        if cfg.dim == 3:
            u0, v0 = S.bubbles_3d(n, cfg.Lx, cfg.seed, idx)
            one = np.ones_like(u0)
            return u0, v0, one, one, {}, "constant", "constant"
        c_kind = cfg.c_classes[idx % len(cfg.c_classes)]
        m_kind = cfg.m_classes[(idx // len(cfg.c_classes)) % len(cfg.m_classes)]
        dev = cfg.sampler == "device"
        if dev:
            cd = F.c_field_device(c_kind, n, cfg.Lx, cfg.seed, idx)
            c = cd.to_host()
            m = F.m_field_device(m_kind, n, cfg.Lx, cfg.seed, idx, c=cd).to_host()
        else:
            c, _ = F.c_field(c_kind, n, cfg.Lx, cfg.seed, idx)
            m, _ = F.m_field(m_kind, n, cfg.Lx, cfg.seed, idx, c=c)
        if self.kind == "nlse":
            if ph == "multi_soliton" and dev:
                psi, prm = S.multi_soliton_device(n, cfg.Lx, cfg.seed, idx)
                psi = psi.to_host()
            elif ph == "multi_soliton":
                psi, prm = S.multi_soliton_2d(n, cfg.Lx, cfg.seed, idx)
            else:
                psi, pk = S.nlse_packets_2d(n, cfg.Lx, cfg.seed, idx)
                prm = {"packets": pk}
            return psi, None, m, c, prm, m_kind, c_kind
        if ph == "kink_field" and dev:
            ud, vd, prm = S.kink_field_device(n, cfg.Lx, cfg.seed, idx)
            u0, v0 = ud.to_host(), vd.to_host()
        elif ph == "kink_field":
            u0, v0, prm = S.kink_field_2d(n, cfg.Lx, cfg.seed, idx)
        elif ph == "kink_ring":
            u0, v0, prm = S.kink_ring_2d(n, cfg.Lx, cfg.seed, idx)
        else:
            u0, v0, g = S.gaussian_sum_2d(n, cfg.Lx, cfg.seed, idx)
            prm = {"gaussians": g}
        return u0, v0, m, c, prm, m_kind, c_kind

    def _write(self, writer, idx, s, u, v, elapsed):
        cfg = self.cfg
        u0, v0, m, c, prm, m_kind = s[:6]
        u = downsample(u, cfg)
        kw = dict(u0=u0, m=m, c=c, u=u, Lx=cfg.Lx, Ly=cfg.Ly, Lz=cfg.Lz, T=cfg.T, nt=cfg.nt,
                  num_snapshots=cfg.num_snapshots, run_id=self.run_id, run_index=idx,
                  phenomenon=self.phenomenon,
                  phenomenon_params=prm, m_type=m_kind, c_type=s[6], elapsed_time=elapsed,
                  problem_type={"kge": "klein_gordon", "nlse": "cubic", "sp4": "phi4"}[self.kind])
        if self.kind != "nlse":
            kw.update(v0=v0, v=downsample(v, cfg))
        writer.submit(self.path(idx), **kw)

    def run(self, log=print):
        """Runs every group.  Host work overlaps the GPU: the next group is sampled and
        the previous one downsampled and handed to the HDF5 writer on worker threads
        while the current group runs (the library releases the interpreter lock inside
        its calls)."""
        from concurrent.futures import ThreadPoolExecutor

        from .h5 import AsyncWriter
        cfg = self.cfg
        t_all = time.perf_counter()
        cells = 0
        n_cells = cfg.nx * cfg.ny * (cfg.nz if cfg.dim == 3 else 1)
        groups = [list(range(g0, min(cfg.num_runs, g0 + cfg.batch)))
                  for g0 in range(0, cfg.num_runs, cfg.batch)]
        self.files = [self.path(i) for g in groups for i in g]

        def write_group(writer, idxs, samples, out, wall):
            for i, s, (u, v) in zip(idxs, samples, out):
                self._write(writer, i, s, u, v, wall / len(idxs))

        with AsyncWriter() as writer, ThreadPoolExecutor(max_workers=2) as pool:
            nxt = pool.submit(lambda g: [self.sample(i) for i in g], groups[0]) if groups else None
            pending = None
            for gi, idxs in enumerate(groups):
                samples = nxt.result()
                if gi + 1 < len(groups):
                    nxt = pool.submit(lambda g: [self.sample(i) for i in g], groups[gi + 1])
                self._first = idxs[0]
                t0 = time.perf_counter()
                out = self._run_group(samples)
                wall = time.perf_counter() - t0
                cells += len(idxs) * n_cells * (cfg.nt - 1)
                if pending is not None:
                    pending.result()
                pending = pool.submit(write_group, writer, idxs, samples, out, wall)
                log(f"[launcher] runs {idxs[0]}..{idxs[-1]}: {wall:.2f} s "
                    f"({len(idxs) * n_cells * (cfg.nt - 1) / wall:.3e} cell-steps/s)")
            if pending is not None:
                pending.result()
            written = writer.drain()
        total = time.perf_counter() - t_all
        return {"files": written, "wall_s": total, "cell_steps": cells, "run_id": self.run_id}

    def _run_group(self, samples):
        cfg = self.cfg
        n = cfg.nx
        dt = cfg.T / cfg.nt
        every = max(cfg.nt // cfg.num_snapshots, 1)
        if self.kind == "sp4":
            # per-trajectory stochastic phi^4 solver (the reference's sp4 executables are
            # per trajectory too); Philox noise keyed by (seed, run)
            from .solvers import SP4SolverDevice
            out = []
            for k, s in enumerate(samples):
                sv = SP4SolverDevice(n, cfg.Lx, cfg.T, cfg.nt, cfg.num_snapshots, s[2], s[3], s[0],
                                     s[1], mode=cfg.integrator or "damped", gamma=cfg.gamma,
                                     sigma=cfg.sigma, seed=cfg.seed, stream_id=self._first + k)
                sv.run()
                out.append(sv.transfer_snapshots())
                sv.close()
            return out
        from .batch import FusedBatch
        method = cfg.integrator or ("strang" if self.kind == "nlse" else "gautschi")
        tile = getattr(self, "_tile", None)  # the first group's autotuned choice, reused
        b = FusedBatch([n] * len(samples), cfg.Lx, dt, equation=self.kind, method=method,
                       dim=cfg.dim, krylov_dim=getattr(self, "_m", cfg.krylov_dim), mu2=cfg.mu2,
                       snapshot_every=every, max_snapshots=cfg.num_snapshots,
                       cells_per_thread=tile["cells_per_thread"] if tile else 0,
                       pass_kernel=tile["pass_kind"] if tile else None,
                       lanes=tile["lanes"] if tile else 0, variant=cfg.lanczos_variant)
        for k, s in enumerate(samples):
            b.set_initial(k, s[2], s[3], s[0], s[1])
        b.run(cfg.nt - 1)  # reference cadence: nt - 1 steps (kg_driver_dev_2d.cpp:144)
        if tile is None and len(samples) == cfg.batch:
            self._tile, self._m = b.tile(), b.info()["krylov_dim"]
        out = [b.snapshots(k) for k in range(len(samples))]
        b.close()
        return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Batched trajectory generation on the GPU "
                                             "(launcher flags of scripts/real_launcher_2d.py)")
    ap.add_argument("--config", help="dataset YAML (config_*.yaml keys)")
    ap.add_argument("--system-type", dest="pde_type", default=None)
    ap.add_argument("--phenomenon", default=None)
    ap.add_argument("--m_type", default=None)
    ap.add_argument("--anisotropy-type", dest="c_type", default=None)
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--Lx", type=float, default=None)
    ap.add_argument("--Ly", type=float, default=None)
    ap.add_argument("--T", type=float, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--snapshots", dest="num_snapshots", type=int, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--output-dir", dest="output_dir", default=None)
    ap.add_argument("--num-runs", dest="num_runs", type=int, default=None)
    ap.add_argument("--dr-x", dest="dr_x", type=int, default=None)
    ap.add_argument("--dr-y", dest="dr_y", type=int, default=None)
    ap.add_argument("--dr-strategy", dest="dr_strategy", default=None,
                    choices=["FFT", "interpolation", "none"])
    ap.add_argument("--integrator", default=None)
    ap.add_argument("--krylov-dim", dest="krylov_dim", type=int, default=None)
    ap.add_argument("--mu2", type=float, default=None)
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--lanczos-variant", dest="lanczos_variant", default=None,
                    choices=("full", "reference"))
    a = ap.parse_args(argv)
    ov = {k: v for k, v in vars(a).items() if k not in ("config", "m_type", "c_type")}
    if a.m_type:
        ov["m_classes"] = [a.m_type]
    if a.c_type:
        ov["c_classes"] = [a.c_type]
    cfg = load_config(a.config, **ov) if a.config else LaunchConfig()
    if not a.config:
        for k, v in ov.items():
            if v is not None:
                setattr(cfg, k, v)
        cfg.ny = cfg.ny if a.ny else cfg.nx
    res = BatchLauncher(cfg).run()
    print(f"[launcher] {res['files']} files in {res['wall_s']:.1f} s, run_id {res['run_id']}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
