"""GPU spectral multislice (HIP phase grating + rocFFT) against the NumPy oracle."""
import ctypes
from decimal import Decimal, getcontext

import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def _sp():
    from wideanglepropagation_amd import spectral
    return spectral


@pytest.mark.parametrize("cdt,rdt,tol", [(torch.complex128, torch.float64, 1e-13),
                                         (torch.complex128, torch.float32, 1e-13),
                                         (torch.complex64, torch.float32, 2e-6),
                                         (torch.complex64, torch.float64, 2e-6)])
def test_transmit_matches_oracle(cdt, rdt, tol):
    sp = _sp()
    rng = np.random.default_rng(0)
    ny, nx, dz, e = 48, 80, 0.03, 200e3
    v = (300.0 * rng.random((ny, nx)) - 20.0).astype(np.float32 if rdt == torch.float32 else np.float64)
    w0 = np.exp(1j * rng.normal(size=(ny, nx))) * rng.random((ny, nx))
    w = torch.tensor(w0, dtype=cdt, device="cuda")
    sp.transmit(w, torch.tensor(v, device="cuda"), dz, e)
    w0c = w0.astype(np.complex64).astype(np.complex128) if cdt == torch.complex64 else w0
    lam = S.energy2wavelength(e)
    ref = w0c * np.exp(2j * np.pi * (S.electron_refractive_index(v.astype(np.float64), e) - 1) * dz / lam)
    got = w.cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < tol


def test_transmit_phase_exact_arithmetic():
    # the kernel forms n - 1 without cancellation: its phase agrees with an exact
    # evaluation to float64 precision relative to the phase itself
    sp = _sp()
    getcontext().prec = 60
    e, dz = 200e3, 0.5
    vs = [1e-4, 0.37, 12.5, 480.0]
    w = torch.ones(len(vs), dtype=torch.complex128, device="cuda")
    sp.transmit(w, torch.tensor(vs, dtype=torch.float64, device="cuda"), dz, e)
    ph = torch.angle(w).cpu().numpy()
    e0 = Decimal(repr(float(S.electron_rest_energy())))
    lam = Decimal(repr(float(S.energy2wavelength(e))))
    pi = Decimal("3.14159265358979323846264338327950288419716939937510")
    for v, p in zip(vs, ph):
        V, E = Decimal(repr(v)), Decimal(repr(e))
        em = E + V
        n = ((2 * em * e0 + em * em) / (2 * E * e0 + E * E)).sqrt()
        exact = float(2 * pi * (n - 1) * Decimal(repr(dz)) / lam)
        assert abs(p - exact) <= 4e-15 * abs(exact) + 1e-18


@pytest.mark.parametrize("kernel", ["fresnel_propagation_kernel",
                                    "angular_spectrum_propagation_kernel"])
def test_kernels_match_oracle(kernel):
    sp = _sp()
    args = (40, 56, (0.04, 0.05), 0.35, 120e3)
    got = getattr(sp, kernel)(*args).cpu().numpy()
    ref = getattr(S, kernel)(*args)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("cdt,tol", [(torch.complex128, 1e-11), (torch.complex64, 2e-5)])
@pytest.mark.parametrize("kernel", ["fresnel_propagation_kernel",
                                    "angular_spectrum_propagation_kernel"])
def test_multislice_matches_oracle(cdt, tol, kernel):
    sp = _sp()
    rng = np.random.default_rng(1)
    N, ny, nx, ps, dz, e = 24, 64, 48, (0.05, 0.05), 0.04, 200e3
    V = 150.0 * rng.random((N, ny, nx)) ** 4
    probe = np.exp(2j * np.pi * rng.random((ny, nx)) * 0.1)
    H = getattr(S, kernel)(ny, nx, ps, dz, e)
    ex, pat, fronts = S.simulate_fresnel_as(V, probe, H, dz, e)
    Hg = getattr(sp, kernel)(ny, nx, ps, dz, e)
    exg, patg, frg = sp.simulate_fresnel_as(torch.tensor(V, device="cuda"),
                                            torch.tensor(probe, dtype=cdt, device="cuda"),
                                            Hg, dz, e)
    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    assert rel(exg.cpu().numpy(), ex) < tol
    assert rel(patg.cpu().numpy(), pat) < 4 * tol
    assert rel(frg.cpu().numpy(), fronts) < tol   # every slice's wavefront
    _, _, none = sp.simulate_fresnel_as(torch.tensor(V, device="cuda"),
                                        torch.tensor(probe, dtype=cdt, device="cuda"), Hg, dz, e,
                                        keep_wavefronts=False)
    assert none is None


def test_compare_exit_waves_figures():
    sp = _sp()
    rng = np.random.default_rng(2)
    a = torch.tensor(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)), device="cuda")
    same = sp.compare_exit_waves(a, 3.0 * a)   # the normalisation removes the scale
    assert same["rel_l2_intensity"] < 1e-14 and same["rel_l2_far_field"] < 1e-14
    assert same["corr_intensity"] == pytest.approx(1.0, abs=1e-14)
    other = sp.compare_exit_waves(a, torch.roll(a, 5, 0))
    assert other["rel_l2_intensity"] > 0.1 and other["corr_intensity"] < 0.9


def test_transmit_errors():
    sp = _sp()
    from wideanglepropagation_amd._lib import lib, DT_C64, DT_F32
    w = torch.ones(8, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        sp.transmit(w.cpu(), torch.zeros(8), 0.1, 200e3)
    with pytest.raises(TypeError):
        sp.transmit(w, torch.zeros(8, dtype=torch.int32, device="cuda"), 0.1, 200e3)
    with pytest.raises(ValueError):
        sp.transmit(w, torch.zeros(9, device="cuda"), 0.1, 200e3)
    L = lib()
    v = torch.zeros(8, device="cuda")
    rc = L.wap_slice_transmit(ctypes.c_void_p(w.data_ptr()), 7, ctypes.c_void_p(v.data_ptr()),
                              DT_F32, 8, 200e3, 511e3, 0.1, 0.025, None)
    assert rc != 0 and b"dtype" in L.wap_last_error()
    rc = L.wap_slice_transmit(ctypes.c_void_p(w.data_ptr()), DT_C64, ctypes.c_void_p(v.data_ptr()),
                              DT_F32, 8, -1.0, 511e3, 0.1, 0.025, None)
    assert rc != 0
    assert L.wap_slice_transmit(None, DT_C64, None, DT_F32, 0, 200e3, 511e3, 0.1, 0.025, None) == 0


def test_fdtd_vs_multislice_quick():
    # the notebook's comparison on its own grid, 8000 steps (the front crosses the
    # 2.1 A from the source to the exit plane in ~5000): runs end to end through the
    # drop-in solver and gives finite figures; the two spectral methods agree closely
    from wideanglepropagation_amd.studies import fdtd_vs_multislice
    r = fdtd_vs_multislice(2.0 / 127, 8000)
    assert r["grid_ext_nz_ny_nx"] == [203, 64, 64] and r["roi_planes"] == 128
    for key in ("fdtd_vs_fresnel", "fdtd_vs_angular_spectrum", "fresnel_vs_angular_spectrum"):
        assert all(np.isfinite(v) for v in r[key].values())
    assert r["fresnel_vs_angular_spectrum"]["corr_intensity"] > 0.99
