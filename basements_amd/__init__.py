"""MI355X-native MLS-MPM soil substep (HIP kernels behind a C ABI).

The compute path is libbm_mpm.so (basements_amd/csrc, built for gfx950 by
basements_amd.build); this package holds its Python host mirror.
"""
from .mpm import (M0, M1, MODE_FAST, MODE_REFERENCE, SHAPE_BOX, SHAPE_CYLINDER, BMError, Body, Collider, MPMSolver,
                  Solver, header_symbols, lib)

__all__ = ["M0", "M1", "MODE_FAST", "MODE_REFERENCE", "SHAPE_BOX", "SHAPE_CYLINDER", "BMError", "Body", "Collider",
           "MPMSolver", "Solver", "header_symbols", "lib"]
