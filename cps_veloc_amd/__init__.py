"""MI355X-native even-odd preconditioned domain-wall / Mobius operator and CG.

The compute path is the HIP library libcps_dwf_gpu.so (C ABI in
include/dwf_gpu.h); this package holds its build script, the ctypes binding
with CPS's DiracOp call surface and the synthetic-input generators.
"""
from . import lattice  # noqa: F401
from .dwf import (CgArg, DeviceVector, DiracOpDwfGpu, DiracOpMobiusGpu,  # noqa: F401
                  DwfGpuError, lib)
