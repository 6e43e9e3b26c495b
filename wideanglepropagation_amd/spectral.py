"""Spectral multislice (Fresnel and angular spectrum) to compare FDTD exit waves with.

A JAX-free restatement, on the GPU, of the spectral propagators of the reference's
``wide_angle_propagation/propagation.py`` that its FDTD notebooks compare exit
waves against (``FDTDmethod/fdtd_3D_baseline_comparison.ipynb``, cells 13-16),
with the same names and argument meaning:

  ``energy2wavelength``                    ``propagation.py:38-50``
  ``electron_refractive_index``            ``propagation.py:66-81``
  ``fresnel_propagation_kernel``           ``propagation.py:139-146``
  ``angular_spectrum_propagation_kernel``  ``propagation.py:149-157``
  ``Propagator``                           ``propagation.py:174-177``
  ``simulate_fresnel_as``                  ``propagation.py:305-352``

Each slice's phase grating (refractive index and ``exp(i 2 pi (n - 1) dz / lambda)``)
is one HIP launch, ``wap_slice_transmit``; the 2D FFTs are rocFFT through
``torch.fft``.  Tensors must be on the GPU: there is no CPU path.  The constants are
ase's defaults (CODATA 2014, ``ase.units``), which the reference imports.

``compare_exit_waves`` gives the comparison figures of the notebook's plots
(cell 15-16: intensities, each normalised by its maximum) as numbers.
"""
import ctypes
import math

import torch

from ._lib import DT_C64, DT_C128, DT_F32, DT_F64, check, lib

# ase.units, CODATA 2014
HPLANCK = 6.626070040e-34   # J s
C_LIGHT = 299792458.0       # m / s
M_E = 9.10938356e-31        # kg
E_CHARGE = 1.6021766208e-19  # C


def electron_rest_energy():
    """m_e c^2 [eV] (``propagation.py:14-22``)."""
    return M_E * C_LIGHT ** 2 / E_CHARGE


def energy2wavelength(energy):
    """Relativistic de Broglie wavelength [A] of an electron of kinetic energy [eV]
    (``propagation.py:38-50``)."""
    return (HPLANCK * C_LIGHT / math.sqrt(energy * (2 * M_E * C_LIGHT ** 2 / E_CHARGE + energy))
            / E_CHARGE * 1.0e10)


def electron_refractive_index(potential, energy):
    """n(V) for the electrostatic potential V [V] (tensor) at kinetic energy [eV]
    (``propagation.py:66-81``), as n - 1 = (n^2 - 1) / (n + 1) with
    n^2 - 1 = V (V + 2 (E0 + E)) / (E (2 E0 + E)) (no cancellation in n - 1)."""
    e0 = electron_rest_energy()
    v = torch.as_tensor(potential).to(torch.float64)
    n2m1 = v * (v + 2.0 * (e0 + energy)) / (energy * (2.0 * e0 + energy))
    return 1.0 + n2m1 / (torch.sqrt(1.0 + n2m1) + 1.0)


def _frequencies(n, m, ps, device):
    """(Fx, Fy) of shape (n, m), indexing 'ij' (``propagation.py:102-107``)."""
    fx = torch.fft.fftfreq(n, ps[0], dtype=torch.float64, device=device)
    fy = torch.fft.fftfreq(m, ps[1], dtype=torch.float64, device=device)
    return torch.meshgrid(fx, fy, indexing="ij")


def fresnel_propagation_kernel(n, m, ps, z, energy, device="cuda"):
    """exp(i 2 pi z / lambda) exp(-i pi lambda z (Fx^2 + Fy^2)), complex128 (n, m)
    (``propagation.py:139-146``); ps = (sampling along axis 0, along axis 1) [A]."""
    lam = energy2wavelength(energy)
    fx, fy = _frequencies(n, m, ps, device)
    ph = (2 * math.pi / lam) * z - math.pi * lam * z * (fx ** 2 + fy ** 2)
    return torch.polar(torch.ones_like(ph), ph)


def angular_spectrum_propagation_kernel(n, m, ps, z, energy, device="cuda"):
    """exp(i 2 pi z kz), kz = sqrt(1/lambda^2 - Fx^2 - Fy^2) taken complex, so that
    evanescent components decay (``propagation.py:149-157``)."""
    lam = energy2wavelength(energy)
    fx, fy = _frequencies(n, m, ps, device)
    kz2 = (1.0 / lam) ** 2 - fx ** 2 - fy ** 2
    kz = torch.sqrt(kz2.to(torch.complex128))
    return torch.exp(1j * 2 * math.pi * z * kz)


def Propagator(u, H):  # noqa: N802 (the reference's name)
    """ifft2(H fft2(u)) (``propagation.py:174-177``)."""
    return torch.fft.ifft2(H * torch.fft.fft2(u))


def _dtype_code(t, kinds):
    codes = {torch.complex64: DT_C64, torch.complex128: DT_C128,
             torch.float32: DT_F32, torch.float64: DT_F64}
    c = codes.get(t.dtype)
    if c is None or c not in kinds:
        raise TypeError(f"unsupported dtype {t.dtype}")
    return c


def transmit(wave, potential_slice, slice_thickness, energy):
    """wave *= exp(i 2 pi (n(V) - 1) dz / lambda) in place, on the GPU (one HIP launch on
    torch's current stream); the phase grating of ``propagation.py:330-335``."""
    if not (wave.is_cuda and potential_slice.is_cuda):
        raise ValueError("transmit runs on the GPU: pass CUDA tensors")
    if wave.shape != potential_slice.shape:
        raise ValueError("wave and potential slice shapes differ")
    if not (wave.is_contiguous() and potential_slice.is_contiguous()):
        raise ValueError("wave and potential slice must be contiguous")
    cw = _dtype_code(wave, (DT_C64, DT_C128))
    cv = _dtype_code(potential_slice, (DT_F32, DT_F64))
    stream = torch.cuda.current_stream(wave.device).cuda_stream
    check(lib().wap_slice_transmit(ctypes.c_void_p(wave.data_ptr()), cw,
                                   ctypes.c_void_p(potential_slice.data_ptr()), cv,
                                   int(wave.numel()), float(energy), electron_rest_energy(),
                                   float(slice_thickness), energy2wavelength(energy),
                                   ctypes.c_void_p(stream)), "wap_slice_transmit")
    return wave


def simulate_fresnel_as(potential, probe, prop_kernel, slice_thickness, energy,
                        keep_wavefronts=True):
    """Multislice through potential (N, ny, nx) [V] from probe (ny, nx) with the
    propagator prop_kernel (Fresnel or angular spectrum) per slice of thickness
    slice_thickness [A] (``propagation.py:305-352``).

    Returns (exit_wave, diffraction_pattern = |fftshift(fft2(exit))|^2, the stacked
    wavefronts after each slice, or None when keep_wavefronts is False).  The wave
    keeps the probe's complex dtype (complex128 as the reference, or complex64)."""
    if not potential.is_cuda:
        raise ValueError("simulate_fresnel_as runs on the GPU: pass CUDA tensors")
    cdt = probe.dtype if probe.dtype in (torch.complex64, torch.complex128) else torch.complex128
    w = probe.to(device=potential.device, dtype=cdt).contiguous().clone()
    H = prop_kernel.to(device=potential.device, dtype=cdt)
    pot = potential.contiguous()
    fronts = (torch.empty((pot.shape[0],) + tuple(w.shape), dtype=cdt, device=w.device)
              if keep_wavefronts else None)
    for i in range(pot.shape[0]):
        transmit(w, pot[i], slice_thickness, energy)   # in place
        w = Propagator(w, H)                           # a new tensor
        if keep_wavefronts:
            fronts[i].copy_(w)   # a copy: the next slice's grating acts on w in place
    det = torch.fft.fftshift(torch.fft.fft2(w))
    pattern = det.real ** 2 + det.imag ** 2
    return w, pattern, fronts


def normalized_intensity(wave):
    """|wave|^2 divided by its maximum (the notebook's cross-section normalisation)."""
    i = wave.abs() ** 2
    m = i.max()
    return i / m if float(m) > 0 else i


def compare_exit_waves(a, b):
    """Figures for two exit waves on the same lateral grid: relative L2 difference of
    their max-normalised intensities, the Pearson correlation of the intensities, and
    the relative L2 difference of their max-normalised far-field intensities."""
    ia, ib = normalized_intensity(a), normalized_intensity(b)
    fa = torch.fft.fftshift(torch.fft.fft2(a)).abs() ** 2
    fb = torch.fft.fftshift(torch.fft.fft2(b)).abs() ** 2
    fa, fb = fa / fa.max(), fb / fb.max()
    x, y = ia - ia.mean(), ib - ib.mean()
    corr = float((x * y).sum() / torch.sqrt((x * x).sum() * (y * y).sum()))
    return {"rel_l2_intensity": float(torch.linalg.vector_norm(ia - ib) / torch.linalg.vector_norm(ib)),
            "corr_intensity": corr,
            "rel_l2_far_field": float(torch.linalg.vector_norm(fa - fb) / torch.linalg.vector_norm(fb))}
