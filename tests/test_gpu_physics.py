"""The reference's GPU physics tests (tests/test_fdtd_3d.py:158-323) through the drop-in API.

Same grid (16 x 16 x 24, extent (-1, 1, -1, 1, -1.5, 1.5)), wavelengths, PML,
source plane, frames and assertions as the reference; the reference gated
them on nvcc, here they run on the HIP backend.  Also: run() snapshots equal
the oracle in the reference-quantised mode, and the fdtd_3d_bin command (the
reference's file contract) gives the same snapshots as the in-process path.
"""
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _solver(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.fdtd_solver import RelativisticSolver3D
    src = tmp_path / "fdtd_3d_params.h"
    shutil.copyfile(os.path.join(ROOT, "wideanglepropagation_amd", "csrc", "fdtd_3d_params.h"),
                    src)
    return RelativisticSolver3D(source_file=str(src), binary_file=str(tmp_path / "fdtd_3d_bin"))


def _run_free_space_case(tmp_path, nx=16, ny=16, nz=24, wavelength=0.1, **kw):
    solver = _solver(tmp_path)
    extent = (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    V_map = np.zeros((nz, ny, nx), dtype=np.float32)
    res = solver.run(V_map, wavelength_angstrom=wavelength, extent=extent, n_frames=2,
                     steps_per_frame=50, snapshot_count=2, pml_thick=0.3, k_source=4,
                     use_angstrom_units=True, **kw)
    return res, extent


def _phase_slope_along_z(field, t, E_sim, extent):
    field_demod = field * np.exp(1j * E_sim * t)
    nz = field_demod.shape[0]
    z = np.linspace(extent[4], extent[5], nz)
    i = field_demod.shape[2] // 2
    j = field_demod.shape[1] // 2
    phase = np.unwrap(np.angle(field_demod[3:nz - 3, j, i]))
    z_fit = z[3:nz - 3]
    slope, intercept = np.polyfit(z_fit, phase, 1)
    return slope, intercept, phase, z_fit


def test_free_space_plane_wave_uniformity(tmp_path):
    res, extent = _run_free_space_case(tmp_path)
    assert res is not None
    t, phi, psi = res["snapshots"][-1]
    intensity = np.abs((phi + 1j * psi)[phi.shape[0] // 2]) ** 2
    assert intensity.mean() > 0
    assert intensity.std() / intensity.mean() < 0.2


def test_symmetric_potential_yields_symmetric_intensity(tmp_path):
    solver = _solver(tmp_path)
    nx, ny, nz = 16, 16, 24
    extent = (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    x = np.linspace(extent[0], extent[1], nx)
    y = np.linspace(extent[2], extent[3], ny)
    z = np.linspace(extent[4], extent[5], nz)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    V = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * 0.3 ** 2))
    V_map = np.transpose(V, (2, 1, 0)).astype(np.float32)
    res = solver.run(V_map, wavelength_angstrom=0.1, extent=extent, n_frames=2,
                     steps_per_frame=50, snapshot_count=2, pml_thick=0.3, k_source=4,
                     use_angstrom_units=True)
    assert res is not None
    t, phi, psi = res["snapshots"][-1]
    intensity = np.abs((phi + 1j * psi)[nz // 2]) ** 2
    assert np.allclose(intensity, intensity[::-1, :], rtol=0.2, atol=1e-3)
    assert np.allclose(intensity, intensity[:, ::-1], rtol=0.2, atol=1e-3)


def test_dirichlet_boundaries_zero(tmp_path):
    res, _ = _run_free_space_case(tmp_path)
    assert res is not None
    _t, phi, psi = res["snapshots"][-1]
    assert np.max(np.abs(phi[[0, -1], :, :]) + np.abs(psi[[0, -1], :, :])) < 1e-3


def test_free_space_phase_matches_wavenumber(tmp_path):
    res, extent = _run_free_space_case(tmp_path, wavelength=0.12)
    assert res is not None
    t, phi, psi = res["snapshots"][-1]
    slope, intercept, phase, z_fit = _phase_slope_along_z(phi + 1j * psi, t, res["E_sim"],
                                                          extent)
    assert np.isfinite(slope)
    assert np.abs(slope) > 1e-3
    residual = phase - (slope * z_fit + intercept)
    r2 = 1.0 - np.sum(residual ** 2) / np.sum((phase - phase.mean()) ** 2)
    assert r2 > 0.8


def test_forward_propagation_phase_slope_positive(tmp_path):
    res, extent = _run_free_space_case(tmp_path, wavelength=0.12)
    assert res is not None
    t, phi, psi = res["snapshots"][-1]
    slope, _, _, _ = _phase_slope_along_z(phi + 1j * psi, t, res["E_sim"], extent)
    assert slope > 0


def test_run_snapshots_match_oracle_reference_quantised(tmp_path):
    from oracle import fdtd3d
    from wideanglepropagation_amd.physics import schrodinger_params
    rng = np.random.default_rng(5)
    nz, ny, nx = 24, 16, 16
    V = (200 * rng.standard_normal((nz, ny, nx))).astype(np.float32)
    res, extent = None, (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    solver = _solver(tmp_path)
    res = solver.run(V, 0.1, extent, n_frames=3, steps_per_frame=7, snapshot_count=2,
                     pml_thick=0.3, k_source=4, precision="f64")
    p = schrodinger_params(0.1, extent, (nz, ny, nx), pml_thick=0.3, k_source=4,
                           reference_quantize=True)
    d = p.oracle_dict()
    st = fdtd3d.zero_state(d)
    outs = []
    for f in range(3):
        st = fdtd3d.run(st, V.astype(np.float64), 7, d, step0=7 * f, f32_time=True)
        outs.append(st)
    assert len(res["snapshots"]) == 2
    for (t, phi, psi), f in zip(res["snapshots"], (1, 2)):
        want = outs[f]
        # the reference stamps a frame with its start time (fdtd_3d.cu:389)
        assert t == np.float32(np.float32(f * 7) * np.float32(p.dt))
        got = phi.astype(np.float64) + 1j * psi.astype(np.float64)
        # float32 snapshots of an f64 run
        assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want)
    assert res["dt"] == np.float32(p.dt) and res["E_sim"] == np.float32(p.e_sim)


def test_cli_backend_matches_in_process(tmp_path):
    from wideanglepropagation_amd.build import CLI_PATH
    if not os.path.exists(CLI_PATH):
        pytest.skip("fdtd_3d_bin not built")
    rng = np.random.default_rng(6)
    V = (300 * rng.standard_normal((24, 16, 16))).astype(np.float32)
    extent = (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    s1 = _solver(tmp_path / "a")
    r1 = s1.run(V, 0.1, extent, n_frames=3, steps_per_frame=5, snapshot_count=2, pml_thick=0.3,
                k_source=4)
    s2 = _solver(tmp_path / "b")
    r2 = s2.run(V, 0.1, extent, n_frames=3, steps_per_frame=5, snapshot_count=2, pml_thick=0.3,
                k_source=4, backend="cli")
    assert r1 is not None and r2 is not None
    assert os.path.exists(os.path.join(s2.working_dir, "snapshots.bin"))
    assert r1["dt"] == r2["dt"] and r1["E_sim"] == r2["E_sim"]
    for (t1, p1, q1), (t2, p2, q2) in zip(r1["snapshots"], r2["snapshots"]):
        assert t1 == t2
        assert np.array_equal(p1, p2) and np.array_equal(q1, q2)


def test_exit_plane_accumulator_matches_notebook_recipe(tmp_path):
    # the notebook recipe (fdtd_3D_baseline_comparison.ipynb:267-290): the last 40 % of the
    # snapshots, demodulated, averaged, divided by sqrt(mean |monitor|^2)
    from oracle import extract
    rng = np.random.default_rng(7)
    V = (100 * rng.standard_normal((24, 16, 16))).astype(np.float32)
    solver = _solver(tmp_path)
    # a random initial state: every plane carries signal from the first step
    psi0 = rng.standard_normal((24, 16, 16)) + 1j * rng.standard_normal((24, 16, 16))
    res = solver.run(V, 0.1, (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5), n_frames=10, steps_per_frame=4,
                     snapshot_count=8, pml_thick=0.3, k_source=4, exit_plane=15,
                     monitor_plane=6, precision="f64", psi0=psi0)
    assert len(res["snapshots"]) == 8
    want, inc = extract.exit_wave(res["snapshots"], float(res["E_sim"]), 15, 6, start_frac=0.6)
    assert inc > 0.1 and res["exit_incident"] == pytest.approx(inc, rel=1e-6)
    assert np.linalg.norm(res["exit_wave"] - want) <= 1e-6 * np.linalg.norm(want)
    # the window matters: the average over every snapshot differs
    every, _ = extract.exit_wave(res["snapshots"], float(res["E_sim"]), 15, 6, start_frac=0.0)
    assert np.linalg.norm(every - want) > 1e-3 * np.linalg.norm(want)


def test_exit_plane_recipe_against_the_fp64_oracle(tmp_path):
    # the same recipe on the oracle's float64 states (snapshot times as the reference
    # stamps them, frame starts on the float32 clock) against the f64 GPU run
    from oracle import extract, fdtd3d
    from wideanglepropagation_amd.physics import schrodinger_params
    rng = np.random.default_rng(8)
    shape, extent = (24, 16, 16), (-1.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    V = (100 * rng.standard_normal(shape)).astype(np.float32)
    psi0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    solver = _solver(tmp_path)
    res = solver.run(V, 0.1, extent, n_frames=10, steps_per_frame=4, snapshot_count=10,
                     pml_thick=0.3, k_source=4, exit_plane=15, monitor_plane=2,
                     precision="f64", no_snapshots=True, psi0=psi0)
    p = schrodinger_params(0.1, extent, shape, pml_thick=0.3, k_source=4,
                           reference_quantize=True)
    d = p.oracle_dict()
    st = psi0
    snaps = []
    for f in range(10):
        st = fdtd3d.run(st, V.astype(np.float64), 4, d, step0=4 * f, f32_time=True)
        t = np.float32(np.float32(4 * f) * np.float32(p.dt))
        snaps.append((t, st.real, st.imag))
    want, inc = extract.exit_wave(snaps, float(np.float32(p.e_sim)), 15, 2, start_frac=0.6)
    assert inc > 0.1 and res["exit_incident"] == pytest.approx(inc, rel=1e-10)
    assert np.linalg.norm(res["exit_wave"] - want) <= 1e-10 * np.linalg.norm(want)
