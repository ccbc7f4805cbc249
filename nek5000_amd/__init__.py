"""MI355X implementation of Nek5000's Helmholtz/Poisson solve path.

libnekamd.so (HIP, gfx950) holds the kernels and the C ABI of
include/nekamd.h; this package is its Python host mirror (solver.Field) and
the synthetic box-mesh generator used by the benchmark and tests.
"""
from ._lib import NekAmdError, load  # noqa: F401

__all__ = ["NekAmdError", "load"]
