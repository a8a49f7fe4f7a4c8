"""MI355X (gfx950) nonlinear angular-spectrum marching.

Drop-in for the per-z march of pinton-lab/angularspectrum_heterogeneous:
``SolverParams`` and ``angular_spectrum_solve(initial_field, params)`` with the
reference's ``(field, pnp, ppp, pI, pIloss, zaxis, pax)`` outputs, running on
hand-written HIP kernels through the C ABI of ``libaship.so``.
"""
from .sources import SourceSlice, generate_phase_screen, make_bowl_source, make_bowl_source_planes, skull_screen
from .transcranial import (build_initial_condition, load_skull_slab, load_sparse_connector, read_nrrd,
                           screens_to_xz_maps, skull_to_phase_screens)

__all__ = [
    "SolverParams", "angular_spectrum_solve", "StabilityError", "MarchPlan",
    "SourceSlice", "make_bowl_source", "make_bowl_source_planes", "generate_phase_screen", "skull_screen",
    "read_nrrd", "load_skull_slab", "skull_to_phase_screens", "screens_to_xz_maps", "load_sparse_connector",
    "build_initial_condition",
]


def __getattr__(name):
    # torch-dependent parts load lazily so the CPU-only pieces import without a GPU
    if name in ("SolverParams", "angular_spectrum_solve", "StabilityError"):
        from . import solver
        return getattr(solver, name)
    if name == "MarchPlan":
        from .plan import MarchPlan
        return MarchPlan
    raise AttributeError(name)
