"""MI355X-native 3D-RISM solute-in-solvent solver (gfx950 HIP + C ABI).

The compute path is ``librism3d.so`` (built by ``python -m bgy3d_amd.build``),
reached through ``bgy3d_amd.rism3d`` (ctypes).  See include/rism3d.h for the
C ABI that replaces the reference's hnc3d_solute_solve.
"""
from . import solutes  # noqa: F401

__all__ = ["solutes", "rism3d"]
