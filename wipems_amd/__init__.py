"""wipems_amd -- MI355X-native FDTD engine for openEMS-style simulations.

Modules:
  engine     Engine (HIP device engine over include/wipems_fdtd.h)
  operator   Operator (host coefficient build over include/wipems_operator.h)
  scenarios  builders of the benchmark and test configurations
  build      in-tree build of the native libraries
"""
from ._native import FdtdError  # noqa: F401

__version__ = "0.1.0"
