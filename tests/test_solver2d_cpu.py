"""Host-side behaviour of the 2D drop-in (``RelativisticSolver``, ``fdtd_solver.py:31-271``).

The reference ships no tests for its 2D class; these follow its 3D unit tests
(``tests/test_fdtd_3d.py:60-155``) on the 2D file contract: ``#define`` rewrite
of NX NY XMIN..YMAX PML_* J_SOURCE, ``potential.bin`` header [nx, ny], the
binary's argv, the ``snapshots.bin`` layout (nx, ny, M, dt, E_sim), and the
legacy ``abs_width`` / ``eta_max`` keywords.
"""
import math
import os
import re
import shutil
import subprocess
import time

import numpy as np
import pytest

from wideanglepropagation_amd.fdtd_solver import (RelativisticSolver, RelativisticSolver3D,
                                                  _au_constants_f32)

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
TEMPLATE = os.path.join(ROOT, "wideanglepropagation_amd", "csrc", "fdtd_params.h")


def _make_solver(tmp_path):
    src = tmp_path / "fdtd_params.h"
    shutil.copyfile(TEMPLATE, src)
    bin_path = tmp_path / "fdtd_open_boundaries"
    bin_path.write_text("")
    return RelativisticSolver(source_file=str(src), binary_file=str(bin_path))


def _write_fake_snapshots(path, nx, ny, M=1, dt=1e-6, E_sim=1.0):
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, M], dtype=np.int32).tobytes())
        f.write(np.array([dt, E_sim], dtype=np.float32).tobytes())
        for i in range(M):
            f.write(np.array([i * dt], dtype=np.float32).tobytes())
            f.write(np.arange(nx * ny, dtype=np.float32).tobytes())
            f.write((-np.arange(nx * ny, dtype=np.float32)).tobytes())


def test_3d_class_inherits_the_2d_class():
    assert issubclass(RelativisticSolver3D, RelativisticSolver)


def test_default_source_file_is_a_private_copy():
    s = RelativisticSolver()
    assert os.path.basename(s.source_file) == "fdtd_params.h"
    assert os.path.dirname(s.source_file) != os.path.dirname(TEMPLATE)
    assert os.path.basename(s.binary_file) == "fdtd_open_boundaries"


def test_update_macros_2d(tmp_path):
    solver = _make_solver(tmp_path)
    changed = solver._update_macros(nx=8, ny=12, xmin=-2.0, xmax=2.0, ymin=-3.0, ymax=3.0,
                                    pml_thick=0.2, pml_sigma=70.0, j_source=5)
    assert changed is True
    content = (tmp_path / "fdtd_params.h").read_text()
    for name, val in (("NX", "8"), ("NY", "12"), ("XMIN", "-2.0000f"), ("YMAX", "3.0000f"),
                      ("PML_THICK", "0.2000f"), ("PML_SIGMA0", "70.0000f"),
                      ("J_SOURCE", "5"), ("PML_POWER", "3.0f")):
        assert re.search(rf"^#define\s+{name}\s+{re.escape(val)}\s*$", content, re.M), name
    assert not re.search(r"^#define\s+NZ\b", content, re.M)
    assert solver._update_macros(8, 12, -2.0, 2.0, -3.0, 3.0, pml_thick=0.2, pml_sigma=70.0,
                                 j_source=5) is False


def test_run_cli_writes_2d_potential_and_argv(tmp_path, monkeypatch):
    solver = _make_solver(tmp_path)
    future = time.time() + 10
    os.utime(solver.binary_file, (future, future))
    _write_fake_snapshots(os.path.join(solver.working_dir, "snapshots.bin"), nx=4, ny=3)
    called = {}

    def fake_run(cmd, check, capture_output, text, cwd):
        called["cmd"] = cmd
        called["cwd"] = cwd
        return subprocess.CompletedProcess(cmd, 0, "", "")

    monkeypatch.setattr(subprocess, "run", fake_run)
    V = np.arange(12, dtype=np.float32).reshape(3, 4)
    res = solver.run(V, wavelength_angstrom=0.05, extent=(-1.0, 1.0, -1.5, 1.5), n_frames=1,
                     steps_per_frame=1, snapshot_count=1, pml_thick=0.3, j_source=1,
                     backend="cli")
    assert res is not None and res["success"]
    assert called["cwd"] == solver.working_dir
    assert called["cmd"][1:5] == ["1", "1", "1", "-0.05"]
    with open(os.path.join(solver.working_dir, "potential.bin"), "rb") as f:
        assert np.frombuffer(f.read(8), dtype=np.int32).tolist() == [4, 3]
        assert np.array_equal(np.frombuffer(f.read(), dtype=np.float32), V.ravel())


def test_load_snapshots_2d_roundtrip(tmp_path):
    solver = _make_solver(tmp_path)
    path = tmp_path / "snapshots.bin"
    _write_fake_snapshots(path, nx=4, ny=3, M=2, dt=2e-6, E_sim=2.5)
    out = solver._load_snapshots(str(path))
    assert out["dt"] == np.float32(2e-6) and out["E_sim"] == np.float32(2.5)
    assert len(out["snapshots"]) == 2
    t1, phi, psi = out["snapshots"][1]
    assert t1 == np.float32(2e-6)
    assert phi.shape == (3, 4) and phi[2, 3] == 11 and psi[1, 0] == -4
    data = path.read_bytes()
    path.write_bytes(data[:-4])
    with pytest.raises(ValueError, match="truncated"):
        solver._load_snapshots(str(path))
    assert solver._load_snapshots(str(tmp_path / "none.bin")) is None


@pytest.mark.parametrize("eta, want", [(60.0, "60.0000f"), (4.0e6, "80.0000f")])
def test_legacy_eta_max_and_abs_width(tmp_path, monkeypatch, eta, want):
    # fdtd_solver.py:170-180: eta_max < 5000 replaces pml_sigma, larger values are ignored
    solver = _make_solver(tmp_path)
    future = time.time() + 10
    os.utime(solver.binary_file, (future, future))
    _write_fake_snapshots(os.path.join(solver.working_dir, "snapshots.bin"), nx=4, ny=3)
    monkeypatch.setattr(subprocess, "run",
                        lambda cmd, **kw: subprocess.CompletedProcess(cmd, 0, "", ""))
    solver.run(np.zeros((3, 4), np.float32), 0.05, (-1, 1, -1, 1), n_frames=1,
               steps_per_frame=1, backend="cli", eta_max=eta, abs_width=0.25)
    content = open(solver.source_file).read()
    assert re.search(rf"^#define\s+PML_SIGMA0\s+{want}", content, re.M)
    assert re.search(r"^#define\s+PML_THICK\s+0\.2500f", content, re.M)


def test_au_constants_follow_the_wavelength():
    # p = 2 pi / lambda_au, E_sim = p^2 / 2, k_beam = sqrt(2 E_sim) = p (fdtd.cu:361-373, 139)
    for lam in (0.05, 0.0251, 0.12):
        p, e, k = _au_constants_f32(lam)
        p_exact = 2 * math.pi / (lam / 0.5291772109)
        assert p == pytest.approx(p_exact, rel=2e-7)
        assert e == pytest.approx(0.5 * p_exact ** 2, rel=4e-7)
        assert k == pytest.approx(p_exact, rel=3e-7)
        assert np.float32(p) == p and np.float32(e) == e and np.float32(k) == k


def test_2d_params_map_onto_the_3d_grid():
    s = RelativisticSolver()
    p = s._params2d((40, 16), (-1.0, 1.0, -2.05, 2.0468), 0.05, 0.4, 80.0, 7, {})
    assert (p.nx, p.ny, p.nz) == (16, 1, 40)
    assert p.extent[4] == float(np.float32(-2.05)) and p.extent[5] == float(np.float32(2.0468))
    assert p.lap_coeff == 0.5 and p.end_planes == "zero" and p.src_k == 7
    assert p.dt == float(np.float32(1e-6)) and p.time_mode == "f32"
    assert p.equation == "schrodinger" and p.precision == "f32"
    d = p.oracle_dict()
    assert d["end_planes"] == "zero" and d["ny"] == 1
    exact = s._params2d((40, 16), (-1.0, 1.0, -2.05, 2.0468), 0.05, 0.4, 80.0, 7,
                        {"reference_quantize": False, "precision": "f64"})
    assert exact.extent[4] == -2.05 and exact.dt == 1e-6 and exact.precision == "f64"
    assert exact.k_beam == pytest.approx(math.sqrt(2 * exact.e_sim), rel=1e-15)
    # one pass per step by default in both precisions; the schedule and chunk keywords
    # reach the grid
    assert p.resolved_schedule() == "schr_1pass" and exact.resolved_schedule() == "schr_1pass"
    st = s._params2d((40, 16), (-1.0, 1.0, -2.05, 2.0468), 0.05, 0.4, 80.0, 7,
                     {"schedule": "stages", "tile_z": 5})
    assert st.resolved_schedule() == "stages" and st.tile_z == 5
