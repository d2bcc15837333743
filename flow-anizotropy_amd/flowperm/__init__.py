"""flowperm -- MI355X lattice-Boltzmann permeability solver.

Drop-in for the solver step of RafalTopolnicki/flow-anizotropy: a voxel mask
in, a 3x3 (D3Q19) or 2x2 (D2Q9) permeability tensor out, computed by HIP
kernels in libflowperm.so (built from ../csrc).
"""
from .structio import StructureError, load_raw, save_raw, load_structure2d, save_gif2d, porosity
from ._lib import FlowpermError

__all__ = ["StructureError", "FlowpermError", "load_raw", "save_raw", "load_structure2d",
           "save_gif2d", "porosity", "solve_permeability3d", "solve_permeability2d",
           "solve_direction3d", "Context"]


def __getattr__(name):
    # the solver module loads the HIP library; import it on first use
    if name in ("solve_permeability3d", "solve_permeability2d", "solve_direction3d", "Context",
                "COLUMNS3D", "COLUMNS2D"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
