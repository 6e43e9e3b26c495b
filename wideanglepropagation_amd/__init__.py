"""MI355X-native 3D RK4 FDTD electron-wave solver.

Drop-in for the ``fdtd_3d.cu`` / ``fdtd_solver.py`` path of
DavidLanders95/WideAnglePropagation: the Python API of ``fdtd_solver``
(``RelativisticSolver3D``, ``calculate_physics_params``) over hand-written
gfx950 HIP kernels reached through the C ABI of ``libwapfdtd.so``
(``include/wap_fdtd.h``).

Modules
  fdtd_solver  the reference's Python API (drop-in)
  physics      constants, ``FDTDParams`` and their builders
  engine       ``FDTD3D``: one device-resident grid (C-ABI wrapper)
  batched      ``BatchedFDTD``: many grids per launch, hipGraph capture / replay
  adjoint      ``ExitWave`` (alias ``FDTDAdjointFunction``), a torch.autograd.Function
               V -> exit wave, and ``exit_wave``
  autotune     tile shapes per size class
  spectral     Fresnel / angular-spectrum multislice for exit-wave comparisons
  workloads    screened-Coulomb slabs and the benchmark grids
  build        hipcc build of ``libwapfdtd.so``
"""
from .physics import (SOL_LIGHT, MASS_PART, BOHR_RADIUS_ANGSTROM,  # noqa: F401
                      calculate_physics_params, FDTDParams, schrodinger_params,
                      klein_gordon_params, beam_constants)
from .fdtd_solver import RelativisticSolver3D  # noqa: F401

__version__ = "0.1.0"


_LAZY = {"FDTD3D": "engine", "BatchedFDTD": "batched", "ExitWave": "adjoint",
         "FDTDAdjointFunction": "adjoint", "exit_wave": "adjoint"}


def __getattr__(name):
    # device-side classes load torch and the HIP library only when asked for
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib
    return getattr(importlib.import_module("." + mod, __name__), name)
