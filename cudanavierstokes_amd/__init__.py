"""MI355X-native explicit compressible Navier-Stokes time advance.

The compute path is libcns.so (hand-written gfx950 HIP kernels behind the
C ABI of include/cns.h); this package is its Python host binding.
"""
from ._lib import CnsConfig, CnsError, load  # noqa: F401
from .solver import (Batch, Config, Solver, boundary_layer, channel, clearSolver, copyField,  # noqa: F401
                     initField, initSolver, mem_info, release_cache, runSimulation, setDevice,
                     setGPUParameters, taylor_green, writeField)

__all__ = ["Batch", "Config", "Solver", "CnsError", "taylor_green", "channel", "boundary_layer", "setDevice",
           "setGPUParameters", "initSolver", "copyField", "runSimulation", "clearSolver", "writeField", "initField", "load",
           "mem_info", "release_cache"]
