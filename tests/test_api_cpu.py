"""Host-side behaviour of the drop-in API and the C ABI (no GPU needed).

The first three tests restate the reference's own unit tests
(``tests/test_fdtd_3d.py:60-155``) against this implementation, with the
same inputs and assertions.
"""
import ctypes
import os
import re
import shutil
import subprocess
import time

import numpy as np
import pytest

from wideanglepropagation_amd import fdtd_solver
from wideanglepropagation_amd.fdtd_solver import RelativisticSolver3D

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
TEMPLATE = os.path.join(ROOT, "wideanglepropagation_amd", "csrc", "fdtd_3d_params.h")
HEADER = os.path.join(ROOT, "include", "wap_fdtd.h")


def _make_solver(tmp_path):
    src = tmp_path / "fdtd_3d_params.h"
    shutil.copyfile(TEMPLATE, src)
    bin_path = tmp_path / "fdtd_3d_bin"
    bin_path.write_text("")
    return RelativisticSolver3D(source_file=str(src), binary_file=str(bin_path))


def _write_fake_snapshots(path, nx, ny, nz, M=1, dt=1e-3, E_sim=1.0):
    vol = nx * ny * nz
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, nz, M], dtype=np.int32).tobytes())
        f.write(np.array([dt, E_sim], dtype=np.float32).tobytes())
        for i in range(M):
            f.write(np.array([i * dt], dtype=np.float32).tobytes())
            f.write(np.zeros(vol, dtype=np.float32).tobytes())
            f.write(np.zeros(vol, dtype=np.float32).tobytes())


def test_update_macros_basic(tmp_path):
    solver = _make_solver(tmp_path)
    changed = solver._update_macros(nx=8, ny=9, nz=10, xmin=-2.0, xmax=2.0, ymin=-3.0, ymax=3.0,
                                    zmin=-4.0, zmax=4.0, pml_thick=0.2, pml_sigma=70.0,
                                    pml_theta=0.7, pml_power=3.0, k_source=5,
                                    wavelength_angstrom=0.05, use_angstrom_units=True)
    assert changed is True
    content = (tmp_path / "fdtd_3d_params.h").read_text()
    assert "#define NX 8" in content
    assert "#define NY 9" in content
    assert "#define NZ 10" in content
    assert "#define XMIN -2.0000f" in content
    assert "#define PML_THICK" in content
    assert "0.2000f" in content
    assert "#define K_SOURCE 5" in content
    # the reference's Angstrom constants at lambda = 0.05 (fdtd_3d.cu:44,48)
    assert "#define LAP_COEFF 3.427719f" in content
    assert "#define P_INF 125.663706f" in content
    # unchanged second call reports False
    assert solver._update_macros(8, 9, 10, -2.0, 2.0, -3.0, 3.0, -4.0, 4.0, pml_thick=0.2,
                                 pml_sigma=70.0, pml_theta=0.7, pml_power=3.0, k_source=5,
                                 wavelength_angstrom=0.05, use_angstrom_units=True) is False


def test_run_writes_potential_and_invokes_binary(tmp_path, monkeypatch):
    solver = _make_solver(tmp_path)
    solver.compile = lambda: None
    future = time.time() + 10
    os.utime(solver.binary_file, (future, future))
    snap_path = os.path.join(solver.working_dir, "snapshots.bin")
    _write_fake_snapshots(snap_path, nx=4, ny=3, nz=2, M=1, dt=1e-3, E_sim=1.0)
    called = {}

    def fake_run(cmd, check, capture_output, text, cwd):
        called["cmd"] = cmd
        called["cwd"] = cwd
        return subprocess.CompletedProcess(cmd, 0, "", "")

    monkeypatch.setattr(subprocess, "run", fake_run)
    V_map = np.zeros((2, 3, 4), dtype=np.float32)
    res = solver.run(V_map, wavelength_angstrom=0.05, extent=(-1.0, 1.0, -1.0, 1.0, -1.5, 1.5),
                     n_frames=1, steps_per_frame=1, snapshot_count=1, pml_thick=0.3, k_source=2,
                     use_angstrom_units=True, backend="cli")
    assert res is not None
    assert "cmd" in called
    assert called["cwd"] == solver.working_dir
    assert called["cmd"][1:5] == ["1", "1", "1", "-0.05"]
    pot_path = os.path.join(solver.working_dir, "potential.bin")
    with open(pot_path, "rb") as f:
        header = np.frombuffer(f.read(12), dtype=np.int32)
        assert header.tolist() == [4, 3, 2]
        assert len(f.read()) == 2 * 3 * 4 * 4


def test_load_snapshots_roundtrip(tmp_path):
    solver = _make_solver(tmp_path)
    snap_path = tmp_path / "snapshots.bin"
    _write_fake_snapshots(snap_path, nx=4, ny=3, nz=2, M=2, dt=2e-3, E_sim=2.5)
    out = solver._load_snapshots(str(snap_path))
    assert out["dt"] == np.float32(2e-3)
    assert out["E_sim"] == np.float32(2.5)
    assert len(out["snapshots"]) == 2
    t0, phi0, psi0 = out["snapshots"][0]
    assert phi0.shape == (2, 3, 4)
    assert psi0.shape == (2, 3, 4)


# ---------------------------------------------------------------- errors
def test_missing_source_file_raises(tmp_path):
    s = RelativisticSolver3D(source_file=str(tmp_path / "nope.h"))
    with pytest.raises(FileNotFoundError):
        s._update_macros(4, 4, 4, -1, 1, -1, 1, -1, 1)


def test_missing_or_truncated_snapshots(tmp_path):
    solver = _make_solver(tmp_path)
    assert solver._load_snapshots(str(tmp_path / "none.bin")) is None
    p = tmp_path / "short.bin"
    _write_fake_snapshots(p, 4, 3, 2, M=2)
    data = p.read_bytes()
    p.write_bytes(data[:-10])
    with pytest.raises(ValueError, match="truncated"):
        solver._load_snapshots(str(p))


def test_cli_failure_returns_none(tmp_path, monkeypatch):
    solver = _make_solver(tmp_path)
    future = time.time() + 10
    os.utime(solver.binary_file, (future, future))

    def failing(cmd, check, capture_output, text, cwd):
        raise subprocess.CalledProcessError(1, cmd, "out", "err")

    monkeypatch.setattr(subprocess, "run", failing)
    res = solver.run(np.zeros((2, 3, 4)), 0.05, (-1, 1, -1, 1, -1, 1), n_frames=1,
                     steps_per_frame=1, backend="cli")
    assert res is None


def test_default_source_file_is_private_copy():
    a = RelativisticSolver3D()
    b = RelativisticSolver3D()
    assert a.source_file != b.source_file
    assert not a.source_file.startswith(os.path.dirname(fdtd_solver.__file__))
    with open(a.source_file) as f:
        assert "#define NX" in f.read()


def test_reference_import_alias():
    import wide_angle_propagation as wap
    assert wap.RelativisticSolver3D is RelativisticSolver3D
    _, p, e = wap.calculate_physics_params(0.05)
    assert p == pytest.approx(66.49836952842307, rel=1e-15)
    from wide_angle_propagation.fdtd_solver import SOL_LIGHT, MASS_PART, BOHR_RADIUS_ANGSTROM
    assert (SOL_LIGHT, MASS_PART, BOHR_RADIUS_ANGSTROM) == (137.035999, 1.0, 0.5291772109)


def test_quantized_params_match_oracle():
    from oracle import physics as ophys
    from wideanglepropagation_amd.physics import schrodinger_params
    ext = (-1.0, 1.0, -1.0, 1.0, -1.6, 1.5811)
    p = schrodinger_params(0.05, ext, (203, 64, 64), pml_thick=0.4, k_source=31,
                           reference_quantize=True)
    q = ophys.quantized_constants(0.05, ext, pml_thick=0.4)
    assert p.lap_coeff == q["A"] and p.e_sim == q["E_sim"] and p.k_beam == q["k_beam"]
    assert p.extent == q["extent"] and p.dt == q["dt"] and p.pml_thick == q["pml_thick"]
    assert p.time_mode == "f32"


# ----------------------------------------------------------------- C ABI
def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(wap_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from wideanglepropagation_amd.build import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    from wideanglepropagation_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(names)
    assert lib.wap_abi_version() == 1


def test_desc_layout_matches_header(tmp_path):
    """Compile a probe against the header with gcc and compare offsets with ctypes."""
    from wideanglepropagation_amd._lib import WapDesc
    fields = [f for f, _ in WapDesc._fields_]
    src = tmp_path / "probe.c"
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(wap_desc, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "wap_fdtd.h"\n'
                   'int main(void){' + body + ' printf("size %zu\\n", sizeof(wap_desc));'
                   ' return 0;}\n')
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = dict(line.split() for line in out if line)
    for f in fields:
        assert int(got[f]) == getattr(WapDesc, f).offset, f
    assert int(got["size"]) == ctypes.sizeof(WapDesc)


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from wideanglepropagation_amd.engine import FDTD3D
    from wideanglepropagation_amd.physics import schrodinger_params
    p = schrodinger_params(0.05, (-1, 1, -1, 1, -1, 1), (8, 4, 4))
    with pytest.raises(RuntimeError, match="GPU"):
        FDTD3D(p)
    # the drop-in API reports the failure the reference way: print + None
    res = RelativisticSolver3D().run(np.zeros((8, 4, 4)), 0.05, (-1, 1, -1, 1, -1, 1),
                                     n_frames=1, steps_per_frame=1)
    assert res is None


def test_package_exports_device_classes_lazily():
    import wideanglepropagation_amd as w
    from wideanglepropagation_amd.adjoint import ExitWave, exit_wave
    from wideanglepropagation_amd.batched import BatchedFDTD
    from wideanglepropagation_amd.engine import FDTD3D
    import torch
    assert w.FDTDAdjointFunction is ExitWave and w.ExitWave is ExitWave
    assert issubclass(w.FDTDAdjointFunction, torch.autograd.Function)
    assert w.exit_wave is exit_wave and w.BatchedFDTD is BatchedFDTD and w.FDTD3D is FDTD3D
    with pytest.raises(AttributeError):
        w.not_a_name


def test_adjoint_segment_plan():
    # the segments tile [0, T) with the full-length one last; the plan is the longest
    # segment within the budget (brute force over every length)
    from wideanglepropagation_amd.adjoint import plan_segment, segment_starts
    for T, S in ((1500, 39), (1500, 118), (10, 3), (10, 10), (10, 25), (7, 1)):
        st = segment_starts(T, S)
        ends = st[1:] + [T]
        assert st[0] == 0 and all(b > a for a, b in zip(st, ends))
        assert all(b - a <= S for a, b in zip(st, ends))
        assert ends[-1] - st[-1] == min(S, T)
        assert len(st) == -(-T // S)
    slot, ck, fixed = 1 << 30, 1 << 30, 5 << 30
    for T in (1, 2, 17, 100, 1500):
        for budget in (0, 8 << 30, 40 << 30, 130 << 30, 2000 << 30):
            lo = max(1, -(-int(T ** 0.5 * 1e6) // 1000000))
            S = plan_segment(T, slot, ck, fixed, budget, lo)
            fits = [s for s in range(lo, T + 1)
                    if s * slot + (-(-T // s) - 1) * ck + fixed <= budget]
            assert S == (max(fits) if fits else lo), (T, budget, S)
    # 128 x 128 x 256 complex64 KG over 1500 steps fits whole in 140 GB: nothing recomputed
    v = 128 * 128 * 256
    assert plan_segment(1500, 16 * v, 16 * v, 40 * v, 140 << 30, 39) == 1500


def test_exit_wave_rejects_no_steps():
    # checked before any device work, so it runs without a GPU
    import torch
    from wideanglepropagation_amd.adjoint import exit_wave
    from wideanglepropagation_amd.physics import klein_gordon_params
    p = klein_gordon_params(0.05, (-0.4, 0.4, -0.4, 0.4, -0.5, 0.5), (12, 8, 8))
    with pytest.raises(ValueError, match="nsteps"):
        exit_wave(torch.zeros((12, 8, 8), dtype=torch.float32), p, 0, 6)


def test_f32_clock_rejects_a_time_origin():
    # the reference's float32 clock (fdtd_3d.cu:365) has no t0: a handle on it refuses one
    # before any device work, so source, demodulation and adjoint seeds share one clock
    from wideanglepropagation_amd.engine import FDTD3D
    from wideanglepropagation_amd.physics import schrodinger_params
    p = schrodinger_params(0.05, (-1, 1, -1, 1, -1, 1), (8, 4, 4), reference_quantize=True)
    assert p.time_mode == "f32"
    with pytest.raises(ValueError, match="t0"):
        FDTD3D(p, t0=1e-3)


def test_exit_wave_rejects_planes_outside_the_grid():
    import torch
    from wideanglepropagation_amd.adjoint import exit_wave
    from wideanglepropagation_amd.physics import klein_gordon_params
    p = klein_gordon_params(0.05, (-0.4, 0.4, -0.4, 0.4, -0.5, 0.5), (12, 8, 8))
    with pytest.raises(ValueError, match="planes"):
        exit_wave(torch.zeros((12, 8, 8), dtype=torch.float32), p, 5, 6, k_monitor=12)


def test_default_samples_are_the_recipe_window():
    # the last 40 % of the frames (start = int(0.6 M)), equal weights, at frame ends
    from wideanglepropagation_amd.adjoint import default_samples
    s = default_samples(1500, 10)
    assert sorted(s) == [1050, 1200, 1350, 1500] and all(w == 0.25 for w in s.values())
    assert default_samples(100, None) == {100: 1.0}


def test_fused_schedules_follow_the_plane_size_not_the_field_size():
    # the fused passes' limit is a chunk window (at least 10 planes) under 2 GiB: a deep
    # grid with 2.12 GiB fields keeps them, a grid whose planes exceed 2^31 / 10 bytes
    # falls back to the 2-pass KG and stage Schrodinger schedules
    from wideanglepropagation_amd.physics import klein_gordon_params, schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent

    def sched(mk, shape, prec="f32"):
        return mk(0.05, centred_extent(shape, 0.05), shape, pml_thick=0.08, k_source=18,
                  precision=prec).resolved_schedule()
    deep = (272, 1024, 1024)
    assert sched(klein_gordon_params, deep) == "kg_fused"
    assert sched(schrodinger_params, deep) == "schr_1pass"
    assert sched(klein_gordon_params, (4096, 1024, 1024), "f64") == "kg_fused"
    wide = (32, 5000, 5440)   # 27.2 M voxels per plane: 217.6 MB in complex64
    assert sched(klein_gordon_params, wide) == "kg_2pass"
    assert sched(schrodinger_params, wide) == "stages"
