"""MI355X-native right-hand side and RK4 time stepping of the EDG discretization
of the Westervelt equation (the hot path of WesterveltEDG.jl).

Host-side table production (``precompute``) is plain numpy; the device API
(``Workspace``, ``edg_rhs_``, ``edg_rhs``, ``rk4_step_``, ``rk4_step``,
``solve``, ``difference_norms``) loads the HIP library libwedg.so on first use
and raises if it is missing. ``edg_westervelt_amd.batch`` advances many
independent instances together (``Batch``, ``solve_many``,
``run_source_many``, the uneven workload ``uneven_instances``).
"""
from .quadrature import JacobiP, GradJacobiP, GaussQCofs
from .coefficients import (Coefficients, paper_coefficients, replace_coefficients,
                           constant_coefficients, element_coefficients)
from .precompute import flux_parameters, GaussianSource, precompute, BOUNDARY_KINDS

_DEVICE_API = ("Workspace", "edg_rhs_", "edg_rhs", "rk4_step_", "rk4_step", "solve",
               "difference_norms", "DimensionMismatch", "WedgError", "mms_errors", "time_step")

__all__ = ["JacobiP", "GradJacobiP", "GaussQCofs", "Coefficients", "paper_coefficients",
           "replace_coefficients", "constant_coefficients", "element_coefficients", "flux_parameters",
           "GaussianSource", "precompute", "BOUNDARY_KINDS", *_DEVICE_API]


def __getattr__(name):
    if name in _DEVICE_API:
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
