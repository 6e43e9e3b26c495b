"""Physical constants and run parameters of the 3D FDTD solver.

Closed forms follow the reference driver ``wide_angle_propagation/fdtd_solver.py``:
  * ``calculate_physics_params`` -- ``fdtd_solver.py:11-29`` (atomic units)
  * Angstrom/eV constants        -- ``fdtd_solver.py:306-342``
  * the decimal quantisation its macro rewrite applies (``fdtd_solver.py:341-367``)
    and the float32 evaluation of E_SIM / k_beam / t in the CUDA program
    (``fdtd_3d.cu:30, 139, 316, 365``), reproduced when ``reference_quantize``.
"""
import math
from dataclasses import dataclass, asdict, replace

import numpy as np

# fdtd_solver.py:7-9
SOL_LIGHT = 137.035999
MASS_PART = 1.0
BOHR_RADIUS_ANGSTROM = 0.5291772109

# fdtd_solver.py:311-313, 332
HBAR_C_EV_A = 1973.2698
REST_ENERGY_EV = 510998.95
HC_EV_A = 12398.4198

# fdtd_3d.cu:30, 59-61, 64-67
DT_DEFAULT = 5e-7
SRC_BAND = 3
SRC_ALPHA = 0.05
PML_THETA_DEFAULT = 0.785
PML_POWER_DEFAULT = 3.0

EQUATIONS = ("schrodinger", "klein_gordon")
SCHEDULES = ("auto", "stages", "kg_2pass", "kg_fused", "schr_fused", "schr_1pass")
PRECISIONS = ("f64", "f32", "f32_kahan")


def calculate_physics_params(wavelength_angstrom):
    """Return (E_total, p_inf, E_sim) in atomic units (``fdtd_solver.py:11-29``)."""
    lambda_au = wavelength_angstrom / BOHR_RADIUS_ANGSTROM
    if lambda_au == 0:
        return 0, 0, 0
    p_inf = (2.0 * np.pi) / lambda_au
    term_pc = p_inf * SOL_LIGHT
    term_mc2 = MASS_PART * SOL_LIGHT ** 2
    e_total = np.sqrt(term_pc ** 2 + term_mc2 ** 2)
    e_sim = 0.5 * p_inf ** 2
    return e_total, p_inf, e_sim


def beam_constants(wavelength_angstrom):
    """Angstrom/eV beam constants (``fdtd_solver.py:332-339``).

    pc = hc/lambda, E_total = sqrt(pc^2 + mc2^2), A = (hbar c)^2/(2 E_total),
    P_INF = 2 pi/lambda, E_sim = A P_INF^2.
    """
    pc = HC_EV_A / wavelength_angstrom
    e_total = math.sqrt(pc * pc + REST_ENERGY_EV * REST_ENERGY_EV)
    a_coef = HBAR_C_EV_A ** 2 / (2.0 * e_total)
    p_inf = 2.0 * math.pi / wavelength_angstrom
    return {"pc": pc, "E_total": e_total, "A": a_coef, "P_INF": p_inf,
            "E_sim": a_coef * p_inf * p_inf, "E_kinetic": e_total - REST_ENERGY_EV}


def wavelength_from_kinetic_energy(e_kin_ev):
    """lambda [A] of an electron of kinetic energy e_kin_ev (inverse of beam_constants)."""
    e_total = e_kin_ev + REST_ENERGY_EV
    pc = math.sqrt(e_total * e_total - REST_ENERGY_EV * REST_ENERGY_EV)
    return HC_EV_A / pc


def _dec(v, fmt):
    return float(np.float32(float(fmt % v)))


@dataclass
class FDTDParams:
    """Everything one grid run needs (the reference's #defines + DT + extensions)."""
    nx: int
    ny: int
    nz: int
    extent: tuple
    lap_coeff: float
    e_sim: float
    k_beam: float
    dt: float = DT_DEFAULT
    equation: str = "schrodinger"
    precision: str = "f32"
    time_mode: str = "exact"          # "exact" | "f32" (fdtd_3d.cu:365)
    hbar_c: float = HBAR_C_EV_A
    rest_energy: float = REST_ENERGY_EV
    pml_thick: float = 0.5
    pml_sigma0: float = 80.0
    pml_theta: float = PML_THETA_DEFAULT
    pml_power: float = PML_POWER_DEFAULT
    kg_gamma_max: float = 0.0
    src_enabled: bool = True
    src_k: int = 20
    src_band: int = SRC_BAND
    src_alpha: float = SRC_ALPHA
    tile_y: int = 0
    tile_z: int = 0
    schedule: str = "auto"            # "auto" | "stages" | "kg_2pass" | "kg_fused" | "schr_fused" | "schr_1pass"
    end_planes: str = "frozen"        # "frozen" (3D reference) | "zero" (2D reference)

    def __post_init__(self):
        if self.equation not in EQUATIONS:
            raise ValueError(f"equation must be one of {EQUATIONS}")
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        if self.schedule not in SCHEDULES:
            raise ValueError(f"schedule must be one of {SCHEDULES}")
        if self.end_planes not in ("frozen", "zero"):
            raise ValueError("end_planes must be 'frozen' or 'zero'")
        if self.time_mode not in ("exact", "f32"):
            raise ValueError("time_mode must be 'exact' or 'f32'")
        self.extent = tuple(float(v) for v in self.extent)
        if len(self.extent) != 6:
            raise ValueError("extent must be (xmin, xmax, ymin, ymax, zmin, zmax)")

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def voxels(self):
        return self.nx * self.ny * self.nz

    def spacing(self):
        x0, x1, y0, y1, z0, z1 = self.extent
        dy = (y1 - y0) / (self.ny - 1.0) if self.ny > 1 else float("inf")
        return ((x1 - x0) / (self.nx - 1.0), dy, (z1 - z0) / (self.nz - 1.0))

    def z_coords(self):
        z0, z1 = self.extent[4], self.extent[5]
        return z0 + np.arange(self.nz) * (z1 - z0) / (self.nz - 1.0)

    def resolved_schedule(self):
        """The pass schedule the library runs for these parameters."""
        # the fused passes address a window of their z chunk and its halo (at least 10
        # planes) with 32-bit offsets
        fits = self.nx * self.ny * (16 if self.precision == "f64" else 8) * 10 < 2 ** 31
        if self.equation != "klein_gordon":
            if self.schedule == "auto":
                if not fits:
                    return "stages"
                # one pass: float32 and float64 (either end rule), float32 + Kahan on 3D
                # grids
                if self.precision in ("f32", "f64") or self.ny > 1:
                    return "schr_1pass"
                return "schr_fused" if self.end_planes == "frozen" else "stages"
            return self.schedule
        if self.schedule == "auto":
            return "kg_fused" if fits else "kg_2pass"
        return self.schedule

    def with_(self, **kw):
        return replace(self, **kw)

    def as_dict(self):
        return asdict(self)

    def oracle_dict(self, src_profile=None):
        """Plain-dict view used by the CPU checker (same numbers, no types)."""
        x0, x1, y0, y1, z0, z1 = self.extent
        return {
            "nx": self.nx, "ny": self.ny, "nz": self.nz,
            "xmin": x0, "xmax": x1, "ymin": y0, "ymax": y1, "zmin": z0, "zmax": z1,
            "dt": self.dt, "A": self.lap_coeff, "e_sim": self.e_sim, "k_beam": self.k_beam,
            "equation": self.equation, "hbar_c": self.hbar_c,
            "rest_energy": self.rest_energy, "pml_thick": self.pml_thick,
            "pml_sigma0": self.pml_sigma0, "pml_theta": self.pml_theta,
            "pml_power": self.pml_power, "kg_gamma_max": self.kg_gamma_max,
            "src_enabled": self.src_enabled, "src_k": self.src_k,
            "src_band": self.src_band, "src_alpha": self.src_alpha,
            "src_profile": src_profile, "end_planes": self.end_planes,
        }


def kg_default_gamma(wavelength_angstrom, pml_thick, pml_power=PML_POWER_DEFAULT,
                     attenuation=1e-6):
    """Sponge strength for the KG absorbing layer.

    With gamma(z) = gamma_max r^p over a layer of thickness T, a wave of group
    velocity v_g = c^2 k / E crossing it once decays by
    exp(-gamma_max T / ((p+1) v_g)); pick gamma_max for the given attenuation.
    """
    c = beam_constants(wavelength_angstrom)
    k = c["P_INF"]
    v_g = HBAR_C_EV_A ** 2 * k / c["E_total"]
    return math.log(1.0 / attenuation) * (pml_power + 1.0) * v_g / pml_thick


def schrodinger_params(wavelength_angstrom, extent, shape, pml_thick=0.5, pml_sigma=80.0,
                       pml_theta=PML_THETA_DEFAULT, pml_power=PML_POWER_DEFAULT, k_source=20,
                       reference_quantize=False, **kw):
    """Run parameters of the reference's Angstrom/eV mode.

    reference_quantize=True reproduces the numbers the compiled reference uses:
    ``%.6f`` A and P_INF, ``%.4f`` extents/PML, ``%.1f`` power
    (``fdtd_solver.py:341-365``) parsed as float; E_sim = A*P*P in float
    (``fdtd_3d.cu:316``); k_beam = sqrtf(E_sim/A) (``:139``); float DT (``:30``)
    and float step times (``:365``).
    """
    nz, ny, nx = shape
    c = beam_constants(wavelength_angstrom)
    if reference_quantize:
        a32 = np.float32(float("%.6f" % c["A"]))
        p32 = np.float32(float("%.6f" % c["P_INF"]))
        e_sim = np.float32(np.float32(a32 * p32) * p32)
        k_beam = np.float32(np.sqrt(np.float32(e_sim / a32)))
        ext = tuple(_dec(v, "%.4f") for v in extent)
        base = dict(lap_coeff=float(a32), e_sim=float(e_sim), k_beam=float(k_beam), extent=ext,
                    pml_thick=_dec(pml_thick, "%.4f"), pml_sigma0=_dec(pml_sigma, "%.4f"),
                    pml_theta=_dec(pml_theta, "%.4f"), pml_power=_dec(pml_power, "%.1f"),
                    dt=float(np.float32(kw.pop("dt", DT_DEFAULT))), time_mode="f32",
                    src_alpha=float(np.float32(SRC_ALPHA)))   # the float literal 0.05f
    else:
        base = dict(lap_coeff=c["A"], e_sim=c["E_sim"], k_beam=c["P_INF"], extent=tuple(extent),
                    pml_thick=pml_thick, pml_sigma0=pml_sigma, pml_theta=pml_theta,
                    pml_power=pml_power, dt=kw.pop("dt", DT_DEFAULT))
    base.update(kw)
    return FDTDParams(nx=nx, ny=ny, nz=nz, equation="schrodinger", src_k=int(k_source), **base)


def klein_gordon_params(wavelength_angstrom, extent, shape, pml_thick=0.5, k_source=20,
                        pml_power=PML_POWER_DEFAULT, kg_gamma_max=None, **kw):
    """First-order Klein-Gordon run: carrier w = E_total, source k = 2 pi / lambda."""
    nz, ny, nx = shape
    c = beam_constants(wavelength_angstrom)
    if kg_gamma_max is None:
        kg_gamma_max = kg_default_gamma(wavelength_angstrom, pml_thick, pml_power)
    return FDTDParams(nx=nx, ny=ny, nz=nz, extent=tuple(extent), equation="klein_gordon",
                      lap_coeff=c["A"], e_sim=c["E_total"], k_beam=c["P_INF"],
                      pml_thick=pml_thick, pml_power=pml_power, kg_gamma_max=kg_gamma_max,
                      src_k=int(k_source), **kw)
