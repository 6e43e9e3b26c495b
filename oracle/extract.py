"""Exit-wave and far-field recipe of the reference notebooks (oracle; test use only).

``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:267-290``:
  k_exit = argmin |z - roi_zmax|, k_monitor = max(0, k_entry - 2);
  for the last 40 % of the snapshots (start = int(0.6 M)) demodulate
  (phi + i psi) e^{+i E_sim t} and average; normalise the exit plane by
  sqrt(mean |psi_monitor|^2) when that mean exceeds 1e-9.
``propagation.py:346-350``: far field = |fftshift(fft2(psi))|^2.
"""
import numpy as np


def demod_phase(e_sim, t):
    """E t of the notebook's phasor (``fdtd_3D_baseline_comparison.ipynb:242, 277``): there
    E_sim and the snapshot time are both numpy float32 (the file's types), so their
    product is rounded to float32; float64 otherwise.  The phasor itself is formed in
    float64 (NumPy 1.x promotes ``1j * float32 scalar`` to complex128)."""
    if isinstance(t, np.float32):
        return float(np.float32(e_sim) * t)
    return float(e_sim) * float(t)


def demodulated_average(snapshots, e_sim, k, start_frac=0.6):
    """Average over snapshots[int(start_frac*M):] of plane k of (phi+i psi) e^{i E t}."""
    m = len(snapshots)
    start = int(start_frac * m)
    acc = None
    count = 0
    for t, phi, psi in snapshots[start:]:
        field = (np.asarray(phi[k], np.float64) + 1j * np.asarray(psi[k], np.float64))
        field = field * np.exp(1j * demod_phase(e_sim, t))
        acc = field if acc is None else acc + field
        count += 1
    if acc is None:
        return None
    return acc / max(1, count)


def exit_wave(snapshots, e_sim, k_exit, k_monitor, start_frac=0.6):
    """Normalised exit wave (``fdtd_3D_baseline_comparison.ipynb:267-290``)."""
    ex = demodulated_average(snapshots, e_sim, k_exit, start_frac)
    mon = demodulated_average(snapshots, e_sim, k_monitor, start_frac)
    inc = float(np.mean(np.abs(mon) ** 2))
    if inc > 1e-9:
        ex = ex / np.sqrt(inc)
    return ex, inc


def far_field(psi_exit):
    """``propagation.py:346-350``."""
    return np.abs(np.fft.fftshift(np.fft.fft2(psi_exit))) ** 2
