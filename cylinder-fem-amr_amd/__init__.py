"""MI355X (gfx950) implementation of cylinder_fem_amr's per-step projection solve.

The product is the HIP kernel library behind include/cfa_hip.h, its C++ host layer behind
include/cfa_host.h and the cylinder_fem_amr program built on them (see Makefile). The Python
module `cfa` binds them with ctypes for in-process runs.
"""
from .cfa import (HISTORY_COLUMNS, LibraryMissing, header_functions, last_history, last_kernels, last_mms,  # noqa: F401
                  last_stats, load, run, uniform_mesh)
