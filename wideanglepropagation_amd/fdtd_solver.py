"""Drop-in replacement of the reference's ``wide_angle_propagation/fdtd_solver.py`` (3D path).

Same names, signatures, defaults, return dictionary and error conventions as
the reference (``fdtd_solver.py:7-29, 274-484``); the stepping runs
in-process on the GPU through ``libwapfdtd.so`` instead of rewriting and
recompiling a CUDA program and exchanging ``potential.bin`` /
``snapshots.bin`` with a child process.

What stays the same
  * ``calculate_physics_params`` and the module constants.
  * ``RelativisticSolver3D(source_file=None, binary_file=None)``,
    ``_update_macros(...) -> bool`` (rewrites ``#define`` lines of
    ``source_file``), ``compile()``, ``run(...) -> dict | None``,
    ``_load_snapshots(path) -> dict | None``.
  * ``run`` returns ``{"snapshots": [(t, phi, psi), ...], "dt", "E_sim",
    "success": True}`` with float32 phi = Re, psi = Im arrays (nz, ny, nx)
    and float32 scalars; failures print and return None.
  * Numbers: by default the constants are quantised exactly as the compiled
    reference sees them (``reference_quantize=True``), the stepper runs in
    float32 like the reference, and snapshot times are the frame *start*
    (``fdtd_3d.cu:389``).

What is new (only through keyword arguments)
  precision ("f32" | "f32_kahan" | "f64"), reference_quantize (bool),
  dt, pml_theta, pml_power (the reference never forwards these),
  equation ("schrodinger" | "klein_gordon"), psi0 / pi0 (initial state),
  source (bool; the reference always injects), source_profile (ny, nx),
  exit_plane (k: the notebooks' exit wave, ``fdtd_3D_baseline_comparison.ipynb:267-290``,
  accumulated on the GPU: plane k demodulated with e^{+i E_sim t} and averaged over
  the saved snapshots from int(exit_start_frac * M) on, default 0.6, i.e. the last
  40 %; with monitor_plane=k_m it is divided by sqrt(mean |average of plane k_m|^2)
  when that exceeds 1e-9; returned as "exit_wave", and "exit_incident"),
  backend ("hip" | "cli": run the
  ``fdtd_3d_bin`` command through the file contract), return_timing,
  schedule ("auto" | "stages" | "schr_fused" | "schr_1pass" | "kg_2pass" | "kg_fused": how a step
  is split into passes; "auto" fuses where the precision allows), tile_y, tile_z.

Deviations (documented): AU mode computes P_INF from the wavelength instead
of leaving a stale value in the source (``fdtd_solver.py:344-349``); the
package file is never modified (the default ``source_file`` is a private
copy of the parameter header).
"""
import os
import re
import shutil
import subprocess
import tempfile
import time

import numpy as np

from .physics import (SOL_LIGHT, MASS_PART, BOHR_RADIUS_ANGSTROM,  # noqa: F401
                      calculate_physics_params, beam_constants, schrodinger_params,
                      klein_gordon_params, FDTDParams, DT_DEFAULT)

__all__ = ["SOL_LIGHT", "MASS_PART", "BOHR_RADIUS_ANGSTROM", "calculate_physics_params",
           "RelativisticSolver", "RelativisticSolver3D"]

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
_PARAMS_TEMPLATE = os.path.join(_PKG_DIR, "csrc", "fdtd_3d_params.h")
_PARAMS_TEMPLATE_2D = os.path.join(_PKG_DIR, "csrc", "fdtd_params.h")
DT_2D = 1e-6   # fdtd.cu DT


def _replace_define(content, name, val_str):
    """``fdtd_solver.py:298-304``: replace the value of ``#define NAME ...``."""
    return re.sub(fr"(#define\s+{name}\s+).*?$", fr"\g<1>{val_str}", content, flags=re.MULTILINE)


def _au_constants_f32(wavelength_angstrom):
    """P_INF, E_sim, k_beam of the 2D program in float (``fdtd.cu`` main, argv[4] < 0:
    lambda_au = lambda / 0.5291772109f, p = 2 pi / lambda_au, E_sim = 0.5 p^2;
    inject_source: k_beam = sqrtf(2 E_sim))."""
    f = np.float32
    lam_au = f(f(abs(wavelength_angstrom)) / f(0.5291772109))
    p_inf = f(f(f(2.0) * f(3.14159265359)) / lam_au)
    e_sim = f(f(f(0.5) * p_inf) * p_inf)
    k_beam = f(np.sqrt(f(f(2.0) * e_sim)))
    return float(p_inf), float(e_sim), float(k_beam)


class RelativisticSolver:
    """2D RK4 FDTD solver (``fdtd_solver.py:31-271``, ``fdtd.cu``) on MI355X.

    The 2D grid (ny, nx) is stepped as a 3D grid with one lateral row: y of the
    reference is the propagation axis (PML, end rows that zero the stage
    outputs, ``fdtd.cu:189-198``), x is periodic; atomic units with
    H = -0.5 Lap + V (``fdtd.cu:250-252``) and DT = 1e-6.
    """
    _RUN_LABEL = "Running"

    def __init__(self, source_file=None, binary_file=None):
        if source_file is None:
            self._tmpdir = tempfile.mkdtemp(prefix="wap_fdtd2d_")
            source_file = os.path.join(self._tmpdir, "fdtd_params.h")
            shutil.copyfile(_PARAMS_TEMPLATE_2D, source_file)
        self.source_file = os.path.abspath(source_file)
        if binary_file is None:
            self.binary_file = os.path.join(os.path.dirname(self.source_file),
                                            "fdtd_open_boundaries")
        else:
            self.binary_file = os.path.abspath(binary_file)
        self.working_dir = os.path.dirname(self.source_file) or "."
        self.last_run = None

    def _update_macros(self, nx, ny, xmin, xmax, ymin, ymax,
                       pml_thick=0.5, pml_sigma=80.0, pml_theta=0.785, pml_power=3.0,
                       j_source=120):
        """``fdtd_solver.py:54-99``: rewrite the #define values; True if changed."""
        if not os.path.exists(self.source_file):
            raise FileNotFoundError(f"Source file not found: {self.source_file}")
        with open(self.source_file, "r") as f:
            content = f.read()
        new = content
        new = _replace_define(new, "NX", str(int(nx)))
        new = _replace_define(new, "NY", str(int(ny)))
        new = _replace_define(new, "XMIN", f"{xmin:.4f}f")
        new = _replace_define(new, "XMAX", f"{xmax:.4f}f")
        new = _replace_define(new, "YMIN", f"{ymin:.4f}f")
        new = _replace_define(new, "YMAX", f"{ymax:.4f}f")
        new = _replace_define(new, "PML_THICK", f"{pml_thick:.4f}f")
        new = _replace_define(new, "PML_SIGMA0", f"{pml_sigma:.4f}f")
        new = _replace_define(new, "PML_THETA", f"{pml_theta:.4f}f")
        new = _replace_define(new, "PML_POWER", f"{pml_power:.1f}f")
        new = _replace_define(new, "J_SOURCE", str(int(j_source)))
        if new != content:
            print(f"Configuration changed. Updating {os.path.basename(self.source_file)}...")
            with open(self.source_file, "w") as f:
                f.write(new)
            return True
        return False

    def compile(self):
        """Make the HIP solver library available (``fdtd_solver.py:101-126``)."""
        from .build import LIB_PATH, build
        try:
            if not os.path.exists(LIB_PATH):
                print("Compiling HIP solver...")
                build()
            from ._lib import lib
            lib()
        except Exception as e:  # noqa: BLE001
            print("Compilation FAILED:")
            print(e)
            raise RuntimeError("Build failed")

    def _params2d(self, shape, extent, wavelength_angstrom, pml_thick, pml_sigma, j_source, kw):
        ny, nx = shape
        xmin, xmax, ymin, ymax = extent
        quant = kw.get("reference_quantize", True)
        pml_theta = kw.get("pml_theta", 0.785)
        pml_power = kw.get("pml_power", 3.0)
        if quant:
            p_inf, e_sim, k_beam = _au_constants_f32(wavelength_angstrom)
            q = lambda v, fmt: float(np.float32(float(fmt % v)))  # noqa: E731
            ext = (q(xmin, "%.4f"), q(xmax, "%.4f"), 0.0, 0.0, q(ymin, "%.4f"), q(ymax, "%.4f"))
            base = dict(dt=float(np.float32(kw.get("dt") or DT_2D)), time_mode="f32",
                        pml_thick=q(pml_thick, "%.4f"), pml_sigma0=q(pml_sigma, "%.4f"),
                        pml_theta=q(pml_theta, "%.4f"), pml_power=q(pml_power, "%.1f"),
                        src_alpha=float(np.float32(0.05)))
        else:
            _, p_inf, e_sim = calculate_physics_params(abs(wavelength_angstrom))
            k_beam = float(np.sqrt(2.0 * e_sim))
            ext = (xmin, xmax, 0.0, 0.0, ymin, ymax)
            base = dict(dt=float(kw.get("dt") or DT_2D), pml_thick=pml_thick,
                        pml_sigma0=pml_sigma, pml_theta=pml_theta, pml_power=pml_power)
        return FDTDParams(nx=nx, ny=1, nz=ny, extent=ext, lap_coeff=0.5, e_sim=e_sim,
                          k_beam=k_beam, equation="schrodinger",
                          precision=kw.get("precision", "f32"), end_planes="zero",
                          src_enabled=bool(kw.get("source", True)), src_k=int(j_source),
                          schedule=str(kw.get("schedule", "auto")),
                          tile_z=int(kw.get("tile_z", 0)), **base)

    def run(self, V_map, wavelength_angstrom, extent,
            n_frames=600, steps_per_frame=100, snapshot_count=None,
            pml_thick=0.5, pml_sigma=80.0, j_source=120, **kwargs):
        """Run the 2D simulation (``fdtd_solver.py:128-242``).

        V_map: (NY, NX) effective potential (Hartree); extent: (xmin, xmax, ymin, ymax)
        in Bohr; y is the propagation axis.  Returns the reference's dict with float32
        (ny, nx) phi / psi snapshots, or None on failure.
        """
        ny, nx = V_map.shape
        xmin, xmax, ymin, ymax = extent
        if snapshot_count is None:
            snapshot_count = n_frames
        if "abs_width" in kwargs:
            pml_thick = kwargs["abs_width"]
        if "eta_max" in kwargs and kwargs["eta_max"] < 5000:   # fdtd_solver.py:171-180
            pml_sigma = kwargs["eta_max"]
        self._update_macros(nx, ny, xmin, xmax, ymin, ymax, pml_thick=pml_thick,
                            pml_sigma=pml_sigma, pml_theta=kwargs.get("pml_theta", 0.785),
                            pml_power=kwargs.get("pml_power", 3.0), j_source=j_source)
        if kwargs.get("backend", "hip") == "cli":
            return self._run_cli(V_map, wavelength_angstrom, n_frames, steps_per_frame,
                                 snapshot_count)
        try:
            self.compile()
            params = self._params2d(V_map.shape, extent, wavelength_angstrom, pml_thick,
                                    pml_sigma, j_source, kwargs)
            V3 = np.asarray(V_map, dtype=np.float32).reshape(ny, 1, nx)
            kw = dict(kwargs)
            if kw.get("psi0") is not None:
                kw["psi0"] = np.asarray(kw["psi0"]).reshape(ny, 1, nx)
            out = self._run_inprocess(params, V3, n_frames, steps_per_frame, snapshot_count,
                                      kw)
        except RuntimeError as e:
            print("Simulation Execution Failed:")
            print(e)
            return None
        out["snapshots"] = [(t, phi.reshape(ny, nx), psi.reshape(ny, nx))
                            for t, phi, psi in out["snapshots"]]
        if "exit_wave" in out:
            out["exit_wave"] = out["exit_wave"].reshape(nx)
        if "state" in out:
            st = out["state"]
            out["state"] = (tuple(s.reshape(ny, nx) for s in st) if isinstance(st, tuple)
                            else st.reshape(ny, nx))
        return out

    def _run_inprocess(self, params, V_map, n_frames, steps_per_frame, snapshot_count, kw):
        import torch
        from .engine import FDTD3D

        V32 = np.ascontiguousarray(np.asarray(V_map).astype(np.float32))
        sim = FDTD3D(params, V=V32)
        try:
            if kw.get("psi0") is not None:
                sim.set_state(kw["psi0"], kw.get("pi0"))
            if kw.get("source_profile") is not None:
                sim.set_source_profile(kw["source_profile"])
            M = int(snapshot_count)
            save_start = max(0, n_frames - M)
            saved = 0
            snaps = []
            # the notebooks' exit-wave recipe (fdtd_3D_baseline_comparison.ipynb:267-290) on
            # device accumulators: the demodulated exit plane (and monitor plane) averaged
            # over the snapshots from int(start_frac * M) on, M the snapshots saved
            exit_k = kw.get("exit_plane")
            mon_k = kw.get("monitor_plane")
            acc = mon = None
            n_saved = min(M, n_frames)
            win0 = int(float(kw.get("exit_start_frac", 0.6)) * n_saved)
            n_acc = 0
            if exit_k is not None:
                acc = torch.zeros((params.ny, params.nx), dtype=torch.complex128, device="cuda")
                if mon_k is not None:
                    mon = torch.zeros_like(acc)
            quant = params.time_mode == "f32"
            dt32 = np.float32(params.dt)
            e_sim32 = np.float32(params.e_sim)
            t_run = 0.0
            for frame in range(n_frames):
                torch.cuda.synchronize()
                t_a = time.perf_counter()
                sim.step(steps_per_frame)
                sim.synchronize()
                t_run += time.perf_counter() - t_a
                if frame >= save_start and saved < M:
                    if quant:
                        t_snap = np.float32(np.float32(frame * steps_per_frame) * dt32)
                    else:
                        t_snap = np.float32((frame + 1) * steps_per_frame * params.dt)
                    if acc is not None and saved >= win0:
                        # E_sim t as the notebook forms it from the snapshot's float32
                        # values: a float32 product (fdtd_3D_baseline_comparison.ipynb:242, 277)
                        ph = float(np.float32(e_sim32 * np.float32(t_snap)))
                        sim.accumulate_plane(exit_k, ph, acc)
                        if mon is not None:
                            sim.accumulate_plane(mon_k, ph, mon)
                        n_acc += 1
                    if not kw.get("no_snapshots", False):
                        st = sim.get_state(np.complex64)
                        if isinstance(st, tuple):
                            st = st[0]
                        phi = np.ascontiguousarray(st.real)
                        psi = np.ascontiguousarray(st.imag)
                        phi.setflags(write=False)
                        psi.setflags(write=False)
                        snaps.append((t_snap, phi, psi))
                    saved += 1
            out = {"snapshots": snaps, "dt": dt32, "E_sim": e_sim32, "success": True}
            vox_steps = params.voxels * n_frames * steps_per_frame
            out["elapsed_s"] = t_run
            out["gvoxel_steps_per_s"] = vox_steps / t_run / 1e9 if t_run > 0 else float("nan")
            if acc is not None:
                ex = (acc / max(1, n_acc)).cpu().numpy()
                if mon is not None:
                    # normalised by the incident intensity at the monitor plane when
                    # it exceeds 1e-9 (the notebook's rule)
                    inc = float(np.mean(np.abs((mon / max(1, n_acc)).cpu().numpy()) ** 2))
                    out["exit_incident"] = inc
                    if inc > 1e-9:
                        ex = ex / np.sqrt(inc)
                out["exit_wave"] = ex
            if kw.get("return_state", False):
                out["state"] = sim.get_state(np.complex128)
            print(f"Loaded {len(snaps)} snapshots.")
            self.last_run = out
            return out
        finally:
            sim.close()

    def _run_cli(self, V_map, wavelength_angstrom, n_frames, steps_per_frame, snapshot_count):
        """File-contract path (``fdtd_solver.py:186-242, 405-454``): potential.bin ->
        the solver command -> snapshots.bin, run in ``working_dir``."""
        from .cli import ensure_cli_binary
        source_mtime = os.path.getmtime(self.source_file)
        binary_mtime = os.path.getmtime(self.binary_file) if os.path.exists(self.binary_file) else 0
        if not os.path.exists(self.binary_file) or source_mtime > binary_mtime:
            self.compile()
            ensure_cli_binary(self.binary_file)
        self._write_potential(V_map)
        cmd = [self.binary_file, str(n_frames), str(steps_per_frame), str(snapshot_count),
               str(-abs(wavelength_angstrom)), self.source_file]
        if os.path.exists(self.binary_file):
            os.chmod(self.binary_file, 0o755)
        print(f"{self._RUN_LABEL}: {' '.join([os.path.basename(c) for c in cmd])}")
        try:
            subprocess.run(cmd, check=True, capture_output=True, text=True, cwd=self.working_dir)
        except subprocess.CalledProcessError as e:
            print("Simulation Execution Failed:")
            print(e.stdout)
            print(e.stderr)
            return None
        return self._load_snapshots(os.path.join(self.working_dir, "snapshots.bin"))

    def _write_potential(self, V_map):
        """``potential.bin`` of the 2D program: int32 [nx, ny] + float32 V
        (``fdtd_solver.py:198-201``)."""
        ny, nx = V_map.shape
        pot_file = os.path.join(self.working_dir, "potential.bin")
        with open(pot_file, "wb") as f:
            f.write(np.array([nx, ny], dtype=np.int32).tobytes())
            f.write(np.asarray(V_map).astype(np.float32).tobytes())
        return pot_file

    def _load_snapshots(self, filepath):
        """``fdtd_solver.py:244-271``: header int32 nx, ny, M, float32 dt, E_sim; M records
        of float32 t, phi[ny,nx], psi[ny,nx]."""
        if not os.path.exists(filepath):
            print(f"Error: Output file {filepath} not found.")
            return None
        snapshots = []
        with open(filepath, "rb") as f:
            head = f.read(20)
            if len(head) != 20:
                raise ValueError(f"{filepath}: truncated header")
            nx, ny, M = np.frombuffer(head[:12], dtype=np.int32)
            dt, e_sim = np.frombuffer(head[12:], dtype=np.float32)
            vol = int(nx) * int(ny)
            for _ in range(int(M)):
                rec = f.read(4 + 8 * vol)
                if len(rec) != 4 + 8 * vol:
                    raise ValueError(f"{filepath}: truncated snapshot record")
                t = np.frombuffer(rec[:4], dtype=np.float32)[0]
                phi = np.frombuffer(rec[4:4 + 4 * vol], dtype=np.float32).reshape((ny, nx))
                psi = np.frombuffer(rec[4 + 4 * vol:], dtype=np.float32).reshape((ny, nx))
                snapshots.append((t, phi, psi))
        print(f"Loaded {len(snapshots)} snapshots.")
        return {"snapshots": snapshots, "dt": dt, "E_sim": e_sim, "success": True}


class RelativisticSolver3D(RelativisticSolver):
    """3D RK4 FDTD solver (``fdtd_solver.py:274-484``) on MI355X."""
    _RUN_LABEL = "Running 3D"

    def __init__(self, source_file=None, binary_file=None):
        if source_file is None:
            # private copy: the package's own files are never rewritten
            self._tmpdir = tempfile.mkdtemp(prefix="wap_fdtd3d_")
            source_file = os.path.join(self._tmpdir, "fdtd_3d_params.h")
            shutil.copyfile(_PARAMS_TEMPLATE, source_file)
        self.source_file = os.path.abspath(source_file)
        if binary_file is None:
            self.binary_file = os.path.join(os.path.dirname(self.source_file), "fdtd_3d_bin")
        else:
            self.binary_file = os.path.abspath(binary_file)
        self.working_dir = os.path.dirname(self.source_file) or "."
        self.last_run = None

    # ------------------------------------------------------------ macros
    def _update_macros(self, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax,
                       pml_thick=0.5, pml_sigma=80.0, pml_theta=0.785, pml_power=3.0,
                       k_source=20, wavelength_angstrom=None, use_angstrom_units=False):
        """Rewrite the ``#define`` values of ``source_file`` (``fdtd_solver.py:285-374``).

        Returns True if the file changed.  Raises FileNotFoundError when the
        file does not exist.
        """
        if not os.path.exists(self.source_file):
            raise FileNotFoundError(f"Source file not found: {self.source_file}")
        with open(self.source_file, "r") as f:
            content = f.read()
        new = content
        if use_angstrom_units and wavelength_angstrom:
            c = beam_constants(wavelength_angstrom)
            new = _replace_define(new, "LAP_COEFF", f"{c['A']:.6f}f")
            new = _replace_define(new, "P_INF", f"{c['P_INF']:.6f}f")
        else:
            new = _replace_define(new, "LAP_COEFF", "0.5f")
        new = _replace_define(new, "NX", str(int(nx)))
        new = _replace_define(new, "NY", str(int(ny)))
        new = _replace_define(new, "NZ", str(int(nz)))
        new = _replace_define(new, "XMIN", f"{xmin:.4f}f")
        new = _replace_define(new, "XMAX", f"{xmax:.4f}f")
        new = _replace_define(new, "YMIN", f"{ymin:.4f}f")
        new = _replace_define(new, "YMAX", f"{ymax:.4f}f")
        new = _replace_define(new, "ZMIN", f"{zmin:.4f}f")
        new = _replace_define(new, "ZMAX", f"{zmax:.4f}f")
        new = _replace_define(new, "PML_THICK", f"{pml_thick:.4f}f")
        new = _replace_define(new, "PML_SIGMA0", f"{pml_sigma:.4f}f")
        new = _replace_define(new, "PML_THETA", f"{pml_theta:.4f}f")
        new = _replace_define(new, "PML_POWER", f"{pml_power:.1f}f")
        new = _replace_define(new, "K_SOURCE", str(int(k_source)))
        if new != content:
            print(f"Configuration changed. Updating {os.path.basename(self.source_file)}...")
            with open(self.source_file, "w") as f:
                f.write(new)
            return True
        return False

    # --------------------------------------------------------------- run
    def _params(self, shape, extent, wavelength_angstrom, pml_thick, pml_sigma, k_source,
                use_angstrom_units, kw):
        nz, ny, nx = shape
        equation = kw.get("equation", "schrodinger")
        precision = kw.get("precision", "f32")
        pml_theta = kw.get("pml_theta", 0.785)
        pml_power = kw.get("pml_power", 3.0)
        quant = kw.get("reference_quantize", True)
        extra = {"precision": precision, "src_enabled": bool(kw.get("source", True))}
        if kw.get("dt") is not None:
            extra["dt"] = float(kw["dt"])
        for key in ("tile_y", "tile_z"):
            if key in kw:
                extra[key] = int(kw[key])
        if "schedule" in kw:
            extra["schedule"] = str(kw["schedule"])
        if equation == "klein_gordon":
            if not use_angstrom_units:
                raise ValueError("klein_gordon needs use_angstrom_units=True")
            return klein_gordon_params(wavelength_angstrom, extent, shape, pml_thick=pml_thick,
                                       k_source=k_source, pml_power=pml_power,
                                       kg_gamma_max=kw.get("kg_gamma_max"), **extra)
        if use_angstrom_units:
            return schrodinger_params(wavelength_angstrom, extent, shape, pml_thick=pml_thick,
                                      pml_sigma=pml_sigma, pml_theta=pml_theta,
                                      pml_power=pml_power, k_source=k_source,
                                      reference_quantize=quant, **extra)
        # AU mode: LAP_COEFF 0.5 (fdtd_solver.py:344-346); P_INF from the wavelength
        _, p_inf, _ = calculate_physics_params(abs(wavelength_angstrom))
        if quant:
            a32 = np.float32(0.5)
            p32 = np.float32(float("%.6f" % p_inf))
            e_sim = float(np.float32(np.float32(a32 * p32) * p32))
            k_beam = float(np.float32(np.sqrt(np.float32(np.float32(e_sim) / a32))))
            dt = float(np.float32(extra.pop("dt", DT_DEFAULT)))
            return FDTDParams(nx=nx, ny=ny, nz=nz,
                              extent=tuple(float(np.float32(float("%.4f" % v))) for v in extent),
                              lap_coeff=0.5, e_sim=e_sim, k_beam=k_beam, dt=dt, time_mode="f32",
                              pml_thick=float(np.float32(float("%.4f" % pml_thick))),
                              pml_sigma0=float(np.float32(float("%.4f" % pml_sigma))),
                              pml_theta=float(np.float32(float("%.4f" % pml_theta))),
                              pml_power=float(np.float32(float("%.1f" % pml_power))),
                              src_alpha=float(np.float32(0.05)), src_k=int(k_source), **extra)
        return FDTDParams(nx=nx, ny=ny, nz=nz, extent=tuple(extent), lap_coeff=0.5,
                          e_sim=0.5 * p_inf * p_inf, k_beam=p_inf, pml_thick=pml_thick,
                          pml_sigma0=pml_sigma, pml_theta=pml_theta, pml_power=pml_power,
                          src_k=int(k_source), **extra)

    def run(self, V_map, wavelength_angstrom, extent,
            n_frames=600, steps_per_frame=100, snapshot_count=None,
            pml_thick=0.5, pml_sigma=80.0, k_source=20,
            use_angstrom_units=True, **kwargs):
        """Run the 3D FDTD simulation (``fdtd_solver.py:376-454``).

        V_map: (NZ, NY, NX) potential energy, eV (Angstrom mode) or Hartree.
        extent: (xmin, xmax, ymin, ymax, zmin, zmax), Angstrom or Bohr.
        """
        nz, ny, nx = V_map.shape
        xmin, xmax, ymin, ymax, zmin, zmax = extent
        if snapshot_count is None:
            snapshot_count = n_frames
        if "abs_width" in kwargs:
            pml_thick = kwargs["abs_width"]
        self._update_macros(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax,
                            pml_thick=pml_thick, pml_sigma=pml_sigma,
                            pml_theta=kwargs.get("pml_theta", 0.785),
                            pml_power=kwargs.get("pml_power", 3.0), k_source=k_source,
                            wavelength_angstrom=wavelength_angstrom,
                            use_angstrom_units=use_angstrom_units)
        if kwargs.get("backend", "hip") == "cli":
            return self._run_cli(V_map, wavelength_angstrom, n_frames, steps_per_frame,
                                 snapshot_count)
        try:
            self.compile()
            params = self._params(V_map.shape, extent, wavelength_angstrom, pml_thick,
                                  pml_sigma, k_source, use_angstrom_units, kwargs)
            return self._run_inprocess(params, V_map, n_frames, steps_per_frame,
                                       snapshot_count, kwargs)
        except RuntimeError as e:
            print("Simulation Execution Failed:")
            print(e)
            return None

    # ------------------------------------------------------------- files
    def _write_potential(self, V_map):
        """``potential.bin``: int32 [nx, ny, nz] + float32 V (``fdtd_solver.py:416-419``)."""
        nz, ny, nx = V_map.shape
        pot_file = os.path.join(self.working_dir, "potential.bin")
        with open(pot_file, "wb") as f:
            f.write(np.array([nx, ny, nz], dtype=np.int32).tobytes())
            f.write(np.asarray(V_map).astype(np.float32).tobytes())
        return pot_file

    def _load_snapshots(self, filepath):
        """Parse ``snapshots.bin`` (``fdtd_solver.py:456-484``).

        Header int32 nx, ny, nz, M, float32 dt, E_sim; then M records of float32 t,
        phi[nz,ny,nx], psi[nz,ny,nx].  Returns None if the file is missing; a
        truncated file raises ValueError.
        """
        if not os.path.exists(filepath):
            print(f"Error: Output file {filepath} not found.")
            return None
        snapshots = []
        with open(filepath, "rb") as f:
            head = f.read(24)
            if len(head) != 24:
                raise ValueError(f"{filepath}: truncated header")
            nx, ny, nz, M = np.frombuffer(head[:16], dtype=np.int32)
            dt, e_sim = np.frombuffer(head[16:], dtype=np.float32)
            vol = int(nx) * int(ny) * int(nz)
            for _ in range(int(M)):
                rec = f.read(4 + 8 * vol)
                if len(rec) != 4 + 8 * vol:
                    raise ValueError(f"{filepath}: truncated snapshot record")
                t = np.frombuffer(rec[:4], dtype=np.float32)[0]
                phi = np.frombuffer(rec[4:4 + 4 * vol], dtype=np.float32).reshape((nz, ny, nx))
                psi = np.frombuffer(rec[4 + 4 * vol:], dtype=np.float32).reshape((nz, ny, nx))
                snapshots.append((t, phi, psi))
        print(f"Loaded {len(snapshots)} snapshots.")
        return {"snapshots": snapshots, "dt": dt, "E_sim": e_sim, "success": True}
