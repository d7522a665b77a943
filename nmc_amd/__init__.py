"""MI355X-native Monte Carlo photon transport (mcsub / mcxyz) behind a C ABI.

The transport kernels live in libnmc_hip.so (HIP, gfx950); this package is the
Python host side: the reference-format parameter structs (params), the
MetalMC-style runtime object (metal_mc), synthetic workloads (workloads) and
the build step (build).
"""
from . import params
from ._lib import NmcError
from .metal_mc import MetalMC, simulate, validate

__all__ = ["params", "NmcError", "MetalMC", "simulate", "validate"]
