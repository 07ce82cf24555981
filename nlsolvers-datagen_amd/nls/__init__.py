"""nls: MI355X (gfx950) implementation of the nlsolvers-datagen time steppers.

Python side of the C-ABI library libnls.so (include/nls.h).  The classes
mirror the reference's C++ classes and the executables' argv so that the
reference's launchers and studies can call into this package:

  operators   build_anisotropic_laplacian_noflux{,_3d}  (common/include/laplacians.hpp)
  device      DeviceArray, DeviceSpMV                    (device/include/spmv.hpp)
  solvers     MatrixFunctionApplicatorReal/Complex, KGESolverDevice,
              NLSESolverDevice, SP4SolverDevice          (device/include/*.hpp)
  batch       FusedBatch: many trajectories in one fused FP32/FP64 launch
  samplers    initial-condition and phenomenon samplers  (scripts/*_sampler.py)
  fields      coefficient-field classes m(x), c(x)       (scripts/m_fields_2d.py,
                                                          scripts/c_fields_2d.py)
  launcher    dataset launcher: YAML/flags -> batch -> per-run HDF5 files
  h5          per-run HDF5 files in the launchers' layout
  packaging   YAML descriptors and per-class NetCDF-4 files of a generated dataset
  server      continuous-batching stochastic phi^4 serving loop

The executables kg_gautschi_2d_dev & co. with the reference's argv are under
bin/ next to this package.
"""
from ._lib import NLSError, load  # noqa: F401

__all__ = ["NLSError", "load"]
