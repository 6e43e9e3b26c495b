"""Spectral multislice (Fresnel / angular spectrum) in NumPy float64 (oracle; test use only).

Restates, JAX-free, the reference's spectral propagators that its FDTD notebooks
compare exit waves against (``FDTDmethod/fdtd_3D_baseline_comparison.ipynb``,
cells 13-16):
  * ``electron_rest_energy``                 -- ``propagation.py:14-22``
  * ``energy2wavelength``                    -- ``propagation.py:38-50``
  * ``electron_refractive_index``            -- ``propagation.py:66-81`` (direct form)
  * ``electron_refractive_index_taylor``     -- ``propagation.py:84-95``
  * ``get_frequencies``                      -- ``propagation.py:102-107``
  * ``fresnel_propagation_kernel``           -- ``propagation.py:139-146``
  * ``angular_spectrum_propagation_kernel``  -- ``propagation.py:149-157``
  * ``propagator``                           -- ``propagation.py:174-177``
  * ``simulate_fresnel_as``                  -- ``propagation.py:305-352``

The physical constants are ase's defaults (CODATA 2014, ``ase.units``), which the
reference imports; ase itself is not installed here.
"""
import numpy as np

# ase.units, CODATA 2014
HPLANCK = 6.626070040e-34   # J s
C_LIGHT = 299792458.0       # m / s
M_E = 9.10938356e-31        # kg
E_CHARGE = 1.6021766208e-19  # C


def electron_rest_energy():
    """m_e c^2 in eV (``propagation.py:14-22``)."""
    return M_E * C_LIGHT ** 2 / E_CHARGE


def energy2wavelength(energy):
    """Relativistic de Broglie wavelength [A] of kinetic energy [eV] (``propagation.py:38-50``)."""
    return (HPLANCK * C_LIGHT / np.sqrt(energy * (2 * M_E * C_LIGHT ** 2 / E_CHARGE + energy))
            / E_CHARGE * 1.0e10)


def electron_refractive_index(potential, energy):
    """n(V) from the electrostatic potential [V] (``propagation.py:66-81``)."""
    e0 = electron_rest_energy()
    v = -np.asarray(potential, dtype=np.float64)
    em = energy - v
    return np.sqrt((2 * em * e0 + em ** 2) / (2 * energy * e0 + energy ** 2))


def electron_refractive_index_taylor(potential, energy):
    """n ~ 1 + (E + E0) / (E (E + 2 E0)) V (``propagation.py:84-95``)."""
    e0 = electron_rest_energy()
    return 1.0 + (energy + e0) / (energy * (energy + 2 * e0)) * np.asarray(potential)


def get_frequencies(n, m, ps):
    """(Fx, Fy), shape (n, m), indexing 'ij' (``propagation.py:102-107``)."""
    fx = np.fft.fftfreq(n, ps[0])
    fy = np.fft.fftfreq(m, ps[1])
    return np.meshgrid(fx, fy, indexing="ij")


def fresnel_propagation_kernel(n, m, ps, z, energy):
    """exp(i 2 pi z / lambda) exp(-i pi lambda z f^2) (``propagation.py:139-146``)."""
    lam = energy2wavelength(energy)
    fx, fy = get_frequencies(n, m, ps)
    return np.exp(1j * (2 * np.pi / lam) * z) * np.exp(-1j * np.pi * lam * z * (fx ** 2 + fy ** 2))


def angular_spectrum_propagation_kernel(n, m, ps, z, energy):
    """exp(i 2 pi z kz), kz = sqrt(1/lambda^2 - f^2) (complex: evanescent modes decay)
    (``propagation.py:149-157``)."""
    lam = energy2wavelength(energy)
    fx, fy = get_frequencies(n, m, ps)
    kz = np.sqrt(np.asarray((1 / lam) ** 2 - fx ** 2 - fy ** 2, dtype=np.complex128))
    return np.exp(1j * 2 * np.pi * z * kz)


def propagator(u, H):
    """ifft2(H fft2(u)) (``propagation.py:174-177``)."""
    return np.fft.ifft2(H * np.fft.fft2(u))


def simulate_fresnel_as(potential, probe, prop_kernel, slice_thickness, energy):
    """Multislice: per slice, the phase grating exp(i 2 pi (n - 1) dz / lambda), then
    propagation by prop_kernel (``propagation.py:305-352``).  Returns (exit wave,
    |fftshift(fft2(exit))|^2, the stacked wavefronts after each slice)."""
    lam = energy2wavelength(energy)
    w = np.asarray(probe, dtype=np.complex128)
    fronts = []
    for i in range(potential.shape[0]):
        n = electron_refractive_index(potential[i], energy)
        w = w * np.exp(1j * 2 * np.pi * (n - 1) * slice_thickness / lam)
        w = propagator(w, prop_kernel)
        fronts.append(w)
    det = np.fft.fftshift(np.fft.fft2(w))
    return w, det.real ** 2 + det.imag ** 2, np.stack(fronts)
