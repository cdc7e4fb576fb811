"""MI355X-native neighbour-list + bond-order + QEq path of SW_REAXFF.

The device path lives in libswreax.so (HIP for gfx950, C-ABI in
include/swreax.h).  This package holds its Python bindings, the
mini-LAMMPS harness binding and the seeded input builders.
"""
from . import systems  # noqa: F401

__all__ = ["systems", "lammps", "_lib"]
