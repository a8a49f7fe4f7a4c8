"""MI355X (gfx950) nonlinear angular-spectrum marching.

Drop-in for the per-z march of pinton-lab/angularspectrum_heterogeneous:
``SolverParams`` and ``angular_spectrum_solve(initial_field, params)`` with the
reference's ``(field, pnp, ppp, pI, pIloss, zaxis, pax)`` outputs, running on
hand-written HIP kernels through the C ABI of ``libaship.so``.
"""
from .sources import SourceSlice, generate_phase_screen, make_bowl_source_planes, skull_screen

__all__ = [
    "SolverParams", "angular_spectrum_solve", "StabilityError", "MarchPlan",
    "SourceSlice", "make_bowl_source_planes", "generate_phase_screen", "skull_screen",
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
