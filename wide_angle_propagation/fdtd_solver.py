"""Import alias of ``wideanglepropagation_amd.fdtd_solver`` under the reference's module path."""
from wideanglepropagation_amd.fdtd_solver import *  # noqa: F401,F403
from wideanglepropagation_amd.fdtd_solver import (RelativisticSolver, RelativisticSolver3D,  # noqa: F401
                                                  calculate_physics_params, SOL_LIGHT,
                                                  MASS_PART, BOHR_RADIUS_ANGSTROM)
