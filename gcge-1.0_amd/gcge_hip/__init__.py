"""gcge_hip: Python access to the gfx950 GCGE backend (libgcge_hip.so).

The product is the C ABI in include/*.h; this package binds it with ctypes
(``capi``), wraps it for tests and benchmarks (``api``) and provides the
seeded problem generators (``problems``).
"""
from . import capi, problems  # noqa: F401
from .api import (Matrix, MultiVec, Ops, Solver, eig_gradient, eig_gradient_generalized, locality_order, solve,
                  synchronize, write_matrix_file)  # noqa: F401
