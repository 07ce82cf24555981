"""Dataset packaging: YAML descriptors and per-class NetCDF-4 files from per-run HDF5 files.

The reference packages a generated dataset in two steps after the runs:
  * parallel_generate_problem_configs.py:213-265 scans the run files and writes one YAML
    descriptor per problem class (pde_type, boundary_condition, num_trajectories,
    num_snapshots, nt, domain{Lx, Ly[, Lz], T}, grid{nx, ny[, nz]}, m_classes,
    c_classes), named config_<pde>_<dims>D_<shape>_ns<ns>_nt<nt>.yaml; the descriptors are
    what canonicalize_data.py:174-262 and the dataset's users read;
  * aggregate_single_core.py:57-206 groups the run files by (dims, stored shape,
    snapshots, domain, complex) and writes one NetCDF-4 file per group;
  * canonicalize_data.py:39-262 rewrites a class file in the canonical layout
    (member, time, x, y[, z]) with coordinate variables and the descriptor's attributes.
Here they read the files this library writes (nls.h5, the launchers' layout); the
NetCDF-4 files are written by the C ABI (nls_nc_class_*, nls_nc_std_*) through libhdf5
with the NetCDF-4 conventions.  The class keys, file names, variables, chunking, attribute
names and YAML key order are the reference's; the problem type and the coefficient
classes come from the files' metadata (problem_type, focusing/m type, c_type) where
the reference parses directory names, which are used as a fallback (m_<class>,
c_<class> path components, kge / nlse in the path).
"""
import ast
import ctypes as C
import os
from collections import defaultdict
from pathlib import Path

import numpy as np
import yaml

from . import _lib as L
from . import h5
from ._lib import call

_PDE = {"klein_gordon": "Klein-Gordon", "cubic": "Nonlinear Schrodinger",
        "nlse": "Nonlinear Schrodinger", "phi4": "phi4", "sp4": "phi4"}


def _path_part(path, prefix):
    """The part after `prefix` of the first path component that starts with it
    (parallel_generate_problem_configs.py:28-33)."""
    for part in Path(path).parts:
        if part.startswith(prefix):
            return part[len(prefix):]
    return None


def _attr(path, obj, name, default=None):
    try:
        return h5.read_attr(path, obj, name)
    except L.NLSError:
        return default


def read_meta(path):
    """Class-relevant metadata of one run file, or None for a file without a stored
    trajectory (aggregate_single_core.py:19-53, parallel_generate_problem_configs.py:36-
    103)."""
    try:
        shape, cplx = h5.shape(path, "u")
    except L.NLSError:
        return None
    if len(shape) < 3:
        return None
    m = {"path": str(path), "dims": len(shape) - 1, "shape": tuple(shape[1:]),
         "ns": int(shape[0]), "is_complex": cplx}
    for k in ("Lx", "Ly", "Lz"):
        v = _attr(path, "grid", k)
        m[k] = float(v) if v is not None else None
    for k in ("nx", "ny", "nz"):
        v = _attr(path, "grid", k)
        m[k] = int(v) if v is not None else None
    T = _attr(path, "time", "T")
    m["T"] = float(T) if T is not None else None
    nt = _attr(path, "time", "nt")
    m["nt"] = int(nt) if nt is not None else None
    ns = _attr(path, "time", "num_snapshots")
    if ns is not None:
        m["ns"] = int(ns)
    pt = _attr(path, "metadata", "problem_type")
    low = str(path).lower()
    if pt is not None:
        m["pde_type"] = _PDE.get(str(pt).lower(), str(pt))
    elif "kge" in low or "klein_gordon" in low:
        m["pde_type"] = "Klein-Gordon"
    elif "nlse" in low or "schrodinger" in low:
        m["pde_type"] = "Nonlinear Schrodinger"
    else:
        m["pde_type"] = "Unknown"
    bc = str(_attr(path, "metadata", "boundary_condition", "neumann")).lower()
    m["boundary"] = "neumann" if bc in ("noflux", "neumann") else bc
    m["m_type"] = _attr(path, "focusing", "type") or _path_part(path, "m_")
    m["c_type"] = _attr(path, "metadata", "c_type") or _path_part(path, "c_")
    return m


def scan(run_dir):
    """Metadata of every readable run file under run_dir, in sorted path order."""
    out = []
    for f in sorted(Path(run_dir).rglob("*.h5")):
        m = read_meta(str(f))
        if m is not None:
            out.append(m)
    return out


# ---------------------------------------------------------------- YAML descriptors
def config_key(m):
    """(pde_type, nx, ny, nz, Lx, Ly, Lz, nt, ns, T, boundary) of a run
    (parallel_generate_problem_configs.py:139-185)."""
    three = m["dims"] == 3
    return (m["pde_type"], m["nx"], m["ny"], m["nz"] if three else None, m["Lx"], m["Ly"],
            m["Lz"] if three else None, m["nt"], m["ns"], m["T"], m["boundary"])


def descriptors(metas):
    """{file name: descriptor dict} with the reference's keys in its order
    (parallel_generate_problem_configs.py:213-265)."""
    groups = defaultdict(lambda: {"count": 0, "m_classes": set(), "c_classes": set()})
    for m in metas:
        g = groups[config_key(m)]
        g["count"] += 1
        if m.get("m_type"):
            g["m_classes"].add(m["m_type"])
        if m.get("c_type"):
            g["c_classes"].add(m["c_type"])
    out = {}
    for key, data in groups.items():
        pde_type, nx, ny, nz, Lx, Ly, Lz, nt, ns, T, boundary = key
        cfg = {"pde_type": pde_type, "boundary_condition": boundary,
               "num_trajectories": data["count"]}
        if ns is not None:
            cfg["num_snapshots"] = int(ns)
        if nt is not None:
            cfg["nt"] = int(nt)
        domain = {}
        if Lx is not None and Ly is not None:
            domain["Lx"] = [float(-Lx), float(Lx)]
            domain["Ly"] = [float(-Ly), float(Ly)]
        if Lz is not None:
            domain["Lz"] = [float(-Lz), float(Lz)]
        if T is not None:
            domain["T"] = [0.0, float(T)]
        if domain:
            cfg["domain"] = domain
        grid = {}
        if nx is not None:
            grid["nx"] = int(nx)
        if ny is not None:
            grid["ny"] = int(ny)
        if nz is not None:
            grid["nz"] = int(nz)
        if grid:
            cfg["grid"] = grid
        if data["m_classes"]:
            cfg["m_classes"] = sorted(data["m_classes"])
        if data["c_classes"]:
            cfg["c_classes"] = sorted(data["c_classes"])
        dims = 3 if nz is not None else 2
        shape_str = f"{nx}x{ny}" if dims == 2 else f"{nx}x{ny}x{nz}"
        name = f"config_{pde_type.replace(' ', '_')}_{dims}D_{shape_str}_ns{ns}_nt{nt}.yaml"
        out[name] = cfg
    return out


def write_descriptors(run_dir, outdir):
    """Writes the YAML descriptors of the runs under run_dir; returns their paths."""
    import yaml
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for name, cfg in descriptors(scan(run_dir)).items():
        p = os.path.join(outdir, name)
        with open(p, "w") as f:
            yaml.dump(cfg, f, default_flow_style=False, sort_keys=False)
        paths.append(p)
    return paths


# ---------------------------------------------------------------- NetCDF-4 class files
def class_key(m):
    """(dims, shape, ns, (Lx, Ly[, Lz]), is_complex) (aggregate_single_core.py:69-75)."""
    L_ = tuple(v for v in (m["Lx"], m["Ly"], m["Lz"] if m["dims"] == 3 else None)
               if v is not None)
    return (m["dims"], tuple(m["shape"]), m["ns"], L_, m["is_complex"])


def classes(metas):
    out = defaultdict(list)
    for m in metas:
        out[class_key(m)].append(m)
    return out


def class_file_name(key, target_shape=None):
    dims, shape, ns, _, cplx = key
    ts = tuple(target_shape) if target_shape is not None else shape
    return (f"class_{dims}D_shape_{'x'.join(map(str, ts))}_ns_{ns}_"
            f"{'complex' if cplx else 'real'}.nc")


def _resample(a, target, meta):
    """a (..., *shape) onto the target grid with the reference's linear interpolation
    (aggregate_single_core.py:86-120 -> downsampling.py:68-103); 2D only here."""
    if tuple(a.shape[-len(target):]) == tuple(target):
        return a
    if meta["dims"] != 2:
        raise ValueError("resampling of 3D classes is not built")
    from .launcher import downsample_interpolation
    lead = a.ndim == 2
    x = a[None] if lead else a
    y = downsample_interpolation(x, target, meta["Lx"], meta["Ly"])
    return y[0] if lead else y


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def write_class(key, metas, outdir, target_shape=None):
    """One NetCDF-4 file of the class `key` from its runs (aggregate_single_core.py:167-
    206): trajectory t at offset t; a run whose data hold NaN is skipped and leaves its
    slot at the fill value, as the reference's loop does.  Returns (path, skipped)."""
    dims, shape, ns, Ls, cplx = key
    target = tuple(target_shape) if target_shape is not None else tuple(shape)
    path = os.path.join(outdir, class_file_name(key, target_shape))
    Lx, Ly = Ls[0], Ls[1]
    Lz = Ls[2] if len(Ls) == 3 else 0.0
    sh = (C.c_int * dims)(*target)
    hd = C.c_void_p()
    call("nls_nc_class_create", path.encode(), dims, sh, int(ns), len(metas), int(cplx),
         float(Lx), float(Ly), float(Lz), C.byref(hd))
    skipped = 0

    def fp(a):
        return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))

    try:
        for t, m in enumerate(metas):
            p = m["path"]
            u = h5.read(p, "u")
            if np.any(np.isnan(u)):
                skipped += 1
                continue
            if cplx:
                a, b = _f32(u.real), _f32(u.imag)
            else:
                a = _f32(u)
                b = _f32(h5.read(p, "v"))
                if np.any(np.isnan(b)):
                    skipped += 1
                    continue
            a, b = _f32(_resample(a, target, m)), _f32(_resample(b, target, m))
            u0 = h5.read(p, "initial_condition/u0")
            if cplx:
                a0, b0 = _f32(u0.real), _f32(u0.imag)
            else:
                a0, b0 = _f32(u0), _f32(h5.read(p, "initial_condition/v0"))
            a0, b0 = _f32(_resample(a0, target, m)), _f32(_resample(b0, target, m))
            mf = _f32(_resample(_f32(h5.read(p, "focusing/m")), target, m))
            cf = _f32(_resample(_f32(h5.read(p, "c" if cplx else "focusing/c")), target, m))
            call("nls_nc_class_put", hd, t, fp(a), fp(b), fp(a0), fp(b0), fp(mf), fp(cf))
    finally:
        call("nls_nc_class_close", hd)
    return path, skipped


def aggregate(run_dir, outdir, target_shape=None):
    """Writes one NetCDF-4 file per class of the runs under run_dir; returns
    {path: (trajectories, skipped)}."""
    os.makedirs(outdir, exist_ok=True)
    out = {}
    for key, metas in classes(scan(run_dir)).items():
        path, skipped = write_class(key, metas, outdir, target_shape)
        out[path] = (len(metas), skipped)
    return out


def _class_attrs(path):
    """Global attributes of a class file (aggregate_single_core.py:193-201)."""
    def ra(k, default=None):
        try:
            return h5.read_attr(path, "/", k)
        except L.NLSError:
            if default is None:
                raise
            return default
    dims = int(ra("dims"))
    shape = tuple(int(x) for x in ast.literal_eval(str(ra("spatial_shape"))))
    return {"num_trajectories": int(ra("num_trajectories")),
            "num_snapshots": int(ra("num_snapshots")), "spatial_shape": shape, "dims": dims,
            "is_complex": bool(int(ra("is_complex"))), "Lx": float(ra("Lx")),
            "Ly": float(ra("Ly")), "Lz": float(ra("Lz", 0.0)) if dims == 3 else None}


def standard_header(path, cfg):
    """The header of the canonical file canonicalize() writes for the class file `path`
    and the dataset descriptor `cfg`, as canonicalize_data.py's dry run prints it (ncdump
    style): dimensions, variables with their attributes, global attributes."""
    a = _class_attrs(path)
    shape, dims = a["spatial_shape"], a["dims"]
    axes = ("x", "y", "z")[:dims]
    out = [f"netcdf {os.path.splitext(os.path.basename(path))[0]} {{", "dimensions:"]
    for name, size in [("member", a["num_trajectories"]), ("time", a["num_snapshots"])] + \
            list(zip(axes, shape)):
        out.append(f"    {name} = {size} ;")
    out += ["", "variables:", "    int32 member(member) ;", "    float32 time(time) ;"]
    out += [f"    float32 {x}({x}) ;" for x in axes]
    fd = ", ".join(("member", "time") + axes)
    for v in (("u_real", "u_imag") if a["is_complex"] else ("u", "v")):
        out.append(f"    float32 {v}({fd}) ;")
    for v in ("m", "c"):
        out.append(f"    float32 {v}({', '.join(('member',) + axes)}) ;")
    out += ["", "// global attributes:"]
    for k, v in _standard_attrs(a, cfg):
        out.append(f"    :{k} = {v} ;")
    out.append("}")
    return "\n".join(out)


def _standard_attrs(a, cfg):
    attrs = [("pde_type", cfg["pde_type"]), ("boundary_condition", cfg["boundary_condition"]),
             ("num_trajectories", a["num_trajectories"]),
             ("num_snapshots", a["num_snapshots"]), ("nt", cfg.get("nt", a["num_snapshots"])),
             ("spatial_shape", str(a["spatial_shape"])), ("dims", a["dims"]),
             ("is_complex", int(a["is_complex"])), ("Lx", a["Lx"]), ("Ly", a["Ly"])]
    if a["Lz"]:
        attrs.append(("Lz", a["Lz"]))
    attrs.append(("T", cfg["domain"]["T"][1]))
    for k in ("m_classes", "c_classes"):
        if k in cfg:
            attrs.append((k, ", ".join(cfg[k])))
    return attrs


def canonicalize(input_file, config_file, dry_run=False):
    """The canonical layout of a per-class NetCDF file (canonicalize_data.py:39-262): the
    flat u, v (re_u, im_u), m, c of `input_file` reshaped to (member, time, x, y[, z]) and
    (member, x, y[, z]) with coordinate variables and the descriptor's global attributes,
    written to `input_file` with `.nc` replaced by `_standard.nc`.  dry_run: print the
    future header and write nothing.  Returns the output path (None for a dry run)."""
    with open(config_file) as f:
        cfg = yaml.safe_load(f)
    if dry_run:
        print(standard_header(str(input_file), cfg))
        return None
    path = str(input_file)
    a = _class_attrs(path)
    n, ns, shape, dims = a["num_trajectories"], a["num_snapshots"], a["spatial_shape"], a["dims"]
    spat = int(np.prod(shape))
    out = path.replace(".nc", "_standard.nc")
    attrs = dict(_standard_attrs(a, cfg))
    sh = (C.c_int * dims)(*shape)
    hd = C.c_void_p()

    def b(k):
        return attrs[k].encode() if k in attrs else None
    call("nls_nc_std_create", out.encode(), dims, sh, ns, n, int(a["is_complex"]),
         a["Lx"], a["Ly"], a["Lz"] or 0.0, float(attrs["T"]), int(attrs["nt"]),
         b("pde_type"), b("boundary_condition"), b("m_classes"), b("c_classes"), C.byref(hd))

    def fp(x):
        return x.ctypes.data_as(C.POINTER(C.c_float))
    names = ("re_u", "im_u") if a["is_complex"] else ("u", "v")
    bu, bv = np.empty(ns * spat, np.float32), np.empty(ns * spat, np.float32)
    bm, bc = np.empty(spat, np.float32), np.empty(spat, np.float32)

    def rd(name, off, buf):  # one trajectory's slice, as the reference reads it
        call("nls_h5_read_range_f32", path.encode(), name.encode(), off, buf.size, fp(buf))
    try:
        for t in range(n):
            rd(names[0], t * ns * spat, bu)
            rd(names[1], t * ns * spat, bv)
            rd("m", t * spat, bm)
            rd("c", t * spat, bc)
            call("nls_nc_std_put", hd, t, fp(bu), fp(bv), fp(bm), fp(bc))
    finally:
        call("nls_nc_std_close", hd)
    return out


def main(argv=None):
    import argparse
    import sys
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "canonicalize":
        # canonicalize_data.py's command line: <netcdf_file> <config_yaml> [--dry-run]
        cp = argparse.ArgumentParser(description="canonical layout of a class file")
        cp.add_argument("command")
        cp.add_argument("netcdf_file")
        cp.add_argument("config_yaml")
        cp.add_argument("--dry-run", action="store_true")
        c = cp.parse_args(argv)
        p = canonicalize(c.netcdf_file, c.config_yaml, c.dry_run)
        if p:
            print(p)
        return
    ap = argparse.ArgumentParser(description="package generated runs: YAML descriptors "
                                             "(describe), per-class NetCDF-4 files (aggregate), "
                                             "or the canonical layout of a class file "
                                             "(canonicalize <file> <config> [--dry-run])")
    ap.add_argument("command", choices=("describe", "aggregate"))
    ap.add_argument("run_dir")
    ap.add_argument("outdir")
    ap.add_argument("--target-shape", type=int, nargs="+", default=None)
    a = ap.parse_args(argv)
    if a.command == "describe":
        for p in write_descriptors(a.run_dir, a.outdir):
            print(p)
    else:
        for p, (n, skipped) in aggregate(a.run_dir, a.outdir, a.target_shape).items():
            print(f"{p}: {n} trajectories, {skipped} skipped")


if __name__ == "__main__":
    main()
