"""MI355X-native right-hand side and RK4 time stepping for the EDG Westervelt solver."""
from .quadrature import JacobiP, GradJacobiP, GaussQCofs
from .coefficients import (Coefficients, paper_coefficients, replace_coefficients,
                           constant_coefficients)
from .precompute import flux_parameters, GaussianSource, precompute, BOUNDARY_KINDS
