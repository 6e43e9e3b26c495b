"""Closed forms of the reference's physical constants (oracle; test use only).

Restates, in plain FP64 Python:
  * ``calculate_physics_params``   -- ``fdtd_solver.py:11-29`` (atomic units)
  * Angstrom/eV constants          -- ``fdtd_solver.py:306-342``
  * the macro quantisation the reference applies when it writes constants
    into the CUDA source as decimal text (``fdtd_solver.py:341-367``) and the
    float32 evaluation of ``E_SIM`` in ``main`` (``fdtd_3d.cu:314-316``) and
    of ``k_beam`` in ``inject_source_3d`` (``fdtd_3d.cu:139``).
"""
import math

import numpy as np

# fdtd_solver.py:7-9
SOL_LIGHT = 137.035999
MASS_PART = 1.0
BOHR_RADIUS_ANGSTROM = 0.5291772109

# fdtd_solver.py:311-313, 332
HBAR_C_EV_A = 1973.2698
REST_ENERGY_EV = 510998.95
HC_EV_A = 12398.4198

# fdtd_3d.cu:30, 59-61
DT_DEFAULT = 5e-7
SRC_BAND = 3
SRC_ALPHA = 0.05


def calculate_physics_params(wavelength_angstrom):
    """AU closed form (``fdtd_solver.py:11-29``): returns (E_total, p_inf, E_sim)."""
    lambda_au = wavelength_angstrom / BOHR_RADIUS_ANGSTROM
    if lambda_au == 0:
        return 0, 0, 0
    p_inf = (2.0 * math.pi) / lambda_au
    term_pc = p_inf * SOL_LIGHT
    term_mc2 = MASS_PART * SOL_LIGHT ** 2
    e_total = math.sqrt(term_pc ** 2 + term_mc2 ** 2)
    e_sim = 0.5 * p_inf ** 2
    return e_total, p_inf, e_sim


def angstrom_constants(wavelength_angstrom):
    """Angstrom/eV mode (``fdtd_solver.py:332-339``).

    Returns dict with pc, E_total (eV), A = (hbar c)^2 / (2 E_total) (eV A^2),
    P_INF = 2 pi / lambda (1/A) and the exact E_sim = A P_INF^2 (eV).
    """
    pc = HC_EV_A / wavelength_angstrom
    e_total = math.sqrt(pc ** 2 + REST_ENERGY_EV ** 2)
    a_coef = HBAR_C_EV_A ** 2 / (2.0 * e_total)
    p_inf = 2.0 * math.pi / wavelength_angstrom
    return {"pc": pc, "E_total": e_total, "A": a_coef, "P_INF": p_inf,
            "E_sim": a_coef * p_inf * p_inf}


def _f32(x):
    return float(np.float32(x))


def quantized_constants(wavelength_angstrom, extent, pml_thick=0.5,
                        pml_sigma=80.0, pml_theta=0.785, pml_power=3.0):
    """The constants exactly as the compiled reference sees them.

    ``_update_macros`` writes ``%.6f`` for LAP_COEFF/P_INF, ``%.4f`` for the
    extents and PML thickness/sigma/theta and ``%.1f`` for the power
    (``fdtd_solver.py:341-365``); nvcc parses them as float literals.
    ``main`` evaluates E_SIM = LAP_COEFF*P_INF*P_INF in float
    (``fdtd_3d.cu:316``); the source uses k_beam = sqrtf(E_sim/LAP_COEFF)
    (``fdtd_3d.cu:139``); DT is the float literal 0.0000005f (``:30``).
    """
    c = angstrom_constants(wavelength_angstrom)
    a32 = np.float32(float("%.6f" % c["A"]))
    p32 = np.float32(float("%.6f" % c["P_INF"]))
    e_sim32 = np.float32(np.float32(a32 * p32) * p32)
    k_beam32 = np.float32(np.sqrt(np.float32(e_sim32 / a32)))
    ext = tuple(_f32(float("%.4f" % v)) for v in extent)
    return {
        "A": float(a32), "P_INF": float(p32), "E_sim": float(e_sim32),
        "k_beam": float(k_beam32), "extent": ext,
        "pml_thick": _f32(float("%.4f" % pml_thick)),
        "pml_sigma0": _f32(float("%.4f" % pml_sigma)),
        "pml_theta": _f32(float("%.4f" % pml_theta)),
        "pml_power": _f32(float("%.1f" % pml_power)),
        "dt": _f32(DT_DEFAULT),
    }
