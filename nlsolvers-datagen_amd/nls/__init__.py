"""nls: MI355X (gfx950) implementation of the nlsolvers-datagen time steppers.

Python side of the C-ABI library libnls.so (include/nls.h).  The classes
mirror the reference's C++ classes and the executables' argv so that the
reference's launchers and studies can call into this package:

  operators   build_anisotropic_laplacian_noflux{,_3d}  (common/include/laplacians.hpp)
  device      DeviceArray, DeviceSpMV                    (device/include/spmv.hpp)
  solvers     MatrixFunctionApplicatorReal/Complex, KGESolverDevice,
              NLSESolverDevice, SP4SolverDevice          (device/include/*.hpp)
  batch       FusedBatch: many trajectories in one fused FP32/FP64 launch
  drivers     kg_gautschi_2d_dev & co. with the reference's argv
"""
from ._lib import NLSError, load  # noqa: F401

__all__ = ["NLSError", "load"]
