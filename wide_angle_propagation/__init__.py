"""Import alias: ``from wide_angle_propagation import RelativisticSolver3D`` keeps working.

The reference package star-imports ``propagation`` (JAX spectral methods,
not part of this solver) before ``fdtd_solver``; this alias exposes only the
FDTD path, served by ``wideanglepropagation_amd``.
"""
from wideanglepropagation_amd.fdtd_solver import *  # noqa: F401,F403
from wideanglepropagation_amd.fdtd_solver import (RelativisticSolver, RelativisticSolver3D,  # noqa: F401
                                                  calculate_physics_params)

__version__ = "0.1.0"
