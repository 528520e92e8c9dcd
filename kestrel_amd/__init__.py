"""kestrel_amd -- Kestrel's finite-volume morphodynamic time step on MI355X.

The product is the HIP library libkestrel_amd.so behind the C-ABI in
include/kestrel_step.h; this package holds its sources (csrc/), the Fortran
iso_c_binding module (fortran/), the in-tree build, and the Python host side
(config, scenario set-up, ctypes runtime).
"""
from . import config, scenario  # noqa: F401

__all__ = ["config", "scenario", "runtime", "workloads"]
