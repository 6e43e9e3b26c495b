"""Workload drivers: precision study, batched probes, adjoint gradient, spectral comparison,
the API entry points.

``python -m wideanglepropagation_amd.studies precision|batched|adjoint|spectral|api [--quick]``
prints one JSON object per study.  Sizes follow the solver's standard
workloads:

precision  Klein-Gordon 200 keV plane wave through a 400-atom slab
           (seed 0) on 512 x 512 x 256 (dx = dy = 0.05 A, dz = 0.005 A,
           PML 16 planes), 4000 steps, in f32 (fused pass), f32_kahan and f64.
           Exit wave = demodulated average of the exit plane over the last 40 %
           of 40 frames (the notebooks' recipe); far field = |fftshift(fft2)|^2.
           Reports relative L2 of exit wave and far field against f64, the norm
           drift trajectory and throughput of each.
batched    48 instances, sizes from {96^3, 128^3, 160x160x256, 192^3}, beam
           energies U[60, 300] keV (seed 1), convergent Gaussian probes at
           random lateral positions over a shared 200-atom slab; alternating
           Schrodinger / Klein-Gordon, complex64 (the fused schedules), 1000
           steps replayed as one hipGraph of 100 steps (the two equation
           groups on forked streams inside it).  Reports aggregate
           Gvoxel-steps/s and the launch shape the autotuner picked: waves per
           equation group, z chunk per size class.
adjoint    Klein-Gordon 128 x 128 x 256, complex64 fields, f32 V, 1500 steps;
           L = || |fft2(exit)|^2 - I_target ||^2 with I_target from a copy of the
           150-atom slab (seed 2) with positions perturbed by N(0, 0.05 A)
           (seed 3).  Reports backward / forward time with sqrt(T) checkpoints
           (segments of ceil(sqrt(T)) steps, every one but the last recomputed) and
           with the default plan (segments as long as half the free device memory
           holds), and checks dL/dV against central differences of the f64
           forward at 32 voxels drawn uniformly (seed 4) between the source and the
           exit plane; the same loss and step count are also timed on the
           512 x 512 x 256 grid of the precision study (timing only).
spectral   The plane-wave comparison of ``fdtd_3D_baseline_comparison.ipynb``:
           lambda = 0.05 A, 64 x 64 over [-1, 1]^2 A, one Z = 79 atom (screened
           Coulomb, 0.3 A) in the ROI z in [-1, 1] A, FDTD (Schrodinger, drop-in
           RelativisticSolver3D) with a 0.4 A PML and 0.2 A source buffer, 30000
           steps, at the notebook's dz = 2/127 A and at dz = 0.005 A; the exit wave
           against the Fresnel and angular spectrum multislices of the same
           potential.  Also the GPU multislice's slices/s at 512 x 512 x 256.
api        The entry points around the stepper: the GPU atom-potential generator
           (400 atoms on 256 x 256 x 512), the reference's RelativisticSolver3D.run
           on that grid (Schrodinger, host V, 1000 steps, one snapshot) and the 2D
           RelativisticSolver.run on 2048 x 4096 (1000 steps), and the
           file-contract command (backend="cli") against the in-process run of a
           64^3 job with 4 snapshots.
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .engine import FDTD3D
from .physics import (klein_gordon_params, schrodinger_params, wavelength_from_kinetic_energy,
                      beam_constants)
from .workloads import kg_slab, atom_potential, random_slab, centred_extent


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def far_field(e):
    return np.abs(np.fft.fftshift(np.fft.fft2(e))) ** 2


def run_exit_wave(p, V, nsteps, n_frames, k_exit, norm_every=None):
    """Step p/V for nsteps; return (exit wave, norm trajectory, seconds of stepping)."""
    sim = FDTD3D(p, V=V)
    spf = nsteps // n_frames
    start = int(0.6 * n_frames)
    acc = torch.zeros((p.ny, p.nx), dtype=torch.complex128, device="cuda")
    norms = []
    t_step = 0.0
    for f in range(n_frames):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(spf)
        sim.synchronize()
        t_step += time.perf_counter() - t0
        if f >= start:
            t = (f + 1) * spf * p.dt
            sim.accumulate_plane(k_exit, p.e_sim * t, acc)
        if norm_every and (f + 1) % norm_every == 0:
            norms.append(((f + 1) * spf, sim.norm2()))
    ex = (acc / max(1, n_frames - start)).cpu().numpy()
    final_norm = sim.norm2()
    sim.close()
    return ex, norms, final_norm, t_step


def precision_study(quick=False):
    shape = (128, 128, 128) if quick else (256, 512, 512)
    nsteps = 400 if quick else 4000
    n_frames = 40
    out = {"study": "precision", "grid_nz_ny_nx": list(shape), "steps": nsteps, "runs": {}}
    ref = None
    # each precision on its default schedule (the fused pass), then the float32 modes on
    # the 2-pass schedule, which differences the neighbours before scaling them
    runs = [("f64", "auto"), ("f32_kahan", "auto"), ("f32", "auto"),
            ("f32_kahan", "kg_2pass"), ("f32", "kg_2pass")]
    for prec, sched in runs:
        p, V = kg_slab(shape, precision=prec, device="cuda")
        p = p.with_(schedule=sched)
        V = V.to(torch.float32)  # every run sees the same float32-rounded potential
        V = V.to(torch.float64) if prec == "f64" else V
        k_exit = p.nz - 16 - 3
        ex, norms, n_final, secs = run_exit_wave(p, V, nsteps, n_frames, k_exit, norm_every=4)
        del V
        torch.cuda.empty_cache()
        r = {"schedule": p.resolved_schedule(),
             "gvoxel_steps_per_s": p.voxels * nsteps / secs / 1e9,
             "norm2_final": n_final, "norm2_trajectory": norms,
             "exit_wave_norm": float(np.linalg.norm(ex))}
        if prec == "f64":
            ref = (ex, far_field(ex), n_final, dict(norms))
        else:
            r["rel_l2_exit_wave_vs_f64"] = rel_l2(ex, ref[0])
            r["rel_l2_far_field_vs_f64"] = rel_l2(far_field(ex), ref[1])
            r["norm_drift_vs_f64"] = (n_final - ref[2]) / ref[2]
            r["norm_drift_trajectory"] = [(s, (n - ref[3][s]) / ref[3][s]) for s, n in norms]
        key = prec if sched == "auto" else f"{prec}/{sched}"
        out["runs"][key] = r
        print(f"[precision] {key}: " + json.dumps({k: v for k, v in r.items()
                                                    if "trajectory" not in k}), flush=True)
    return out


def probe_profile(shape, extent, x0, y0, lam, width=1.0, focus=-5.0):
    """Convergent Gaussian probe at (x0, y0): Gaussian envelope with a quadratic phase
    focusing at distance ``focus`` beyond the source plane (paraxial)."""
    nz, ny, nx = shape
    x = np.linspace(extent[0], extent[1], nx)
    y = np.linspace(extent[2], extent[3], ny)
    Lx, Ly = nx * (x[1] - x[0]), ny * (y[1] - y[0])
    X = (x[None, :] - x0 + Lx / 2) % Lx - Lx / 2
    Y = (y[:, None] - y0 + Ly / 2) % Ly - Ly / 2
    r2 = X ** 2 + Y ** 2
    k = 2 * math.pi / lam
    return np.exp(-r2 / (2 * width ** 2)) * np.exp(1j * k * r2 / (2 * focus))


def batched_study(quick=False, n_inst=48, nsteps=1000, autotune=True):
    from .batched import BatchedFDTD
    from .autotune import FIXED_WAVES, tune_batch
    rng = np.random.default_rng(1)
    classes = [(96, 96, 96), (128, 128, 128), (256, 160, 160), (192, 192, 192)]
    if quick:
        classes = [(32, 24, 24), (40, 32, 32), (48, 24, 24), (36, 36, 36)]
        n_inst, nsteps = 8, 50
    spacing, spacing_z = 0.05, 0.005
    # shared slab: 200 atoms in the largest lateral cell, wrapped into each grid
    big = max(c[1] for c in classes) * spacing
    pos, Zs, scr = random_slab(200, ((-big / 2, big / 2), (-big / 2, big / 2)), (-0.2, 0.2), 11)
    sims, meta = [], []
    shapes = [classes[i] for i in rng.integers(0, len(classes), n_inst)]
    energies = rng.uniform(60e3, 300e3, n_inst)
    params = []
    for i, (shape, ekin) in enumerate(zip(shapes, energies)):
        eq = "schrodinger" if i % 2 == 0 else "klein_gordon"
        lam = wavelength_from_kinetic_energy(ekin)
        ext = centred_extent(shape, spacing, spacing_z)
        mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
        params.append(mk(lam, ext, shape, pml_thick=16 * spacing_z, k_source=18,
                         precision="f32"))
    tuned, tuned_ms = {}, None
    if autotune:
        # the fused schedules (the float32 default): workgroups of 16 waves; the z chunk
        # per size class, timed on the whole batch
        tuned, tuned_ms, _ = tune_batch(params, tile_zs=(16, 32, 64), tile_ys=(8, 16),
                                        tile_y=16)
        print(f"[batched] tuned z chunks: {tuned_ms:.3f} ms per batched step", file=sys.stderr,
              flush=True)
    for i, (shape, p) in enumerate(zip(shapes, params)):
        eq = p.equation
        ekin = float(energies[i])
        lam = wavelength_from_kinetic_energy(ekin)
        ext = centred_extent(shape, spacing, spacing_z)
        cls = (p.equation, p.precision, p.resolved_schedule(), p.shape)
        if cls in tuned:
            p = p.with_(tile_y=tuned[cls][0], tile_z=tuned[cls][1])
        V = atom_potential(shape, ext, pos, Zs, scr, r_soft=spacing, cutoff=4.0, device="cuda")
        x0, y0 = rng.uniform(ext[0], ext[1]), rng.uniform(ext[2], ext[3])
        prof = probe_profile(shape, ext, x0, y0, lam)
        sims.append(FDTD3D(p, V=V.to(torch.float32), source_profile=prof))
        meta.append({"eq": eq, "shape": list(shape), "E_kin_eV": float(ekin)})
        if (i + 1) % 8 == 0:
            print(f"[batched] {i + 1}/{n_inst} instances set up", file=sys.stderr, flush=True)
        del V
    b = BatchedFDTD(sims, tile_y=4)
    chunk = min(nsteps, 100)
    b.capture(chunk)
    b.replay()          # warm-up
    b.synchronize()
    for s in sims:
        s.set_state(None)
    b.step_index = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    done = 0
    while done < nsteps:
        dev_ms += b.replay(timed=True)
        done += chunk
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    vox = sum(s.params.voxels for s in sims)
    norms = [s.norm2() for s in sims]
    out = {"study": "batched", "instances": n_inst, "steps": done,
           "total_voxels": vox, "gvoxel_steps_per_s": vox * done / wall / 1e9,
           "gvoxel_steps_per_s_device": vox * done / (dev_ms * 1e-3) / 1e9,
           "graph_steps_per_launch": chunk, "launch_groups": b.tiles(),
           "autotuned": {f"{k[0]}/{'x'.join(map(str, k[3]))}":
                         {"waves": FIXED_WAVES.get(k[2], v[0]), "tile_z": v[1]}
                         for k, v in tuned.items()},
           "autotuned_ms_per_batched_step": tuned_ms,
           "all_finite": bool(np.all(np.isfinite(norms))),
           "instances_meta": meta[:6]}
    b.close()
    for s in sims:
        s.close()
    return out


def adjoint_study(quick=False, n_fd=32):
    from .adjoint import exit_wave, far_field_intensity
    from .adjoint import last_plan as adjoint_plan
    shape = (64, 32, 32) if quick else (256, 128, 128)
    nsteps = 150 if quick else 1500
    if quick:
        n_fd = 6
    spacing, spacing_z = 0.05, 0.005
    ext = centred_extent(shape, spacing, spacing_z)
    half = ext[5] - 20 * spacing_z
    pos, Zs, scr = random_slab(150, ((ext[0], ext[1]), (ext[2], ext[3])), (-half, half), 2)
    pert = pos + np.random.default_rng(3).normal(0.0, 0.05, pos.shape)
    p32, V = kg_slab(shape, precision="f32", device="cuda", positions=(pos, Zs, scr))
    # dt = 1e-6 (stable: max |w| dt ~ 0.9): 1500 steps carry the wave front across the
    # 1.1 A between the source and the exit plane (at 5e-7 it would not yet arrive)
    p32 = p32.with_(dt=1e-6)
    _, Vt = kg_slab(shape, precision="f32", device="cuda", positions=(pert, Zs, scr))
    k_exit = p32.nz - 16 - 3
    n_frames = 10
    spf = nsteps // n_frames
    samples = {(f + 1) * spf: 1.0 / 4 for f in range(6, 10)}
    V32 = V.to(torch.float32)
    sqrt_t = math.ceil(math.sqrt(nsteps))
    with torch.no_grad():
        It = far_field_intensity(exit_wave(Vt.to(torch.float32), p32, nsteps, k_exit,
                                           samples=samples))

    def loss_of(Vx, p, **kw):
        E = exit_wave(Vx, p, nsteps, k_exit, samples=samples, **kw)
        return torch.sum((far_field_intensity(E) - It) ** 2)

    with torch.no_grad():
        L0 = loss_of(V32, p32)
    # the cost with sqrt(T) checkpoints (every segment but the last recomputed), and with the
    # default plan (the longest segments half the free device memory holds)
    costs = {}
    for plan_name, kw in (("sqrt_t", {"checkpoint_every": sqrt_t}), ("memory_budget", {})):
        t_fwd, t_fb, Vg = _time_forward_backward(lambda Vx: loss_of(Vx, p32, **kw), V32)
        costs[plan_name] = {"segment_plan": dict(adjoint_plan), "t_forward_s": t_fwd,
                            "t_forward_backward_s": t_fb,
                            "backward_over_forward": (t_fb - t_fwd) / t_fwd}
        print(f"[adjoint] {plan_name}: forward {t_fwd:.3f} s, forward+backward {t_fb:.3f} s",
              file=sys.stderr, flush=True)
    g32 = Vg.grad.double().cpu().numpy()
    # f64 adjoint for the same loss
    p64 = p32.with_(precision="f64")
    Vd = V32.double().clone().requires_grad_(True)
    L64 = loss_of(Vd, p64, checkpoint_every=sqrt_t)
    L64.backward()
    g64 = Vd.grad.cpu().numpy()
    # central differences of the f64 forward at voxels drawn uniformly (seed 4) from the
    # planes the wave crosses between the source and the exit plane; a voxel whose
    # difference cannot resolve 1e-4 (|fd| below 1e4 times its float64 rounding,
    # 4 eps |L| / 2h) is drawn again
    rng = np.random.default_rng(4)
    hstep = 1e-3 * float(V32.abs().max())
    resolvable = 1e4 * 4 * np.finfo(np.float64).eps * abs(L64.item()) / (2 * hstep)
    rows, redrawn = [], 0
    Vbase = V32.double()
    while len(rows) < n_fd and redrawn <= n_fd:
        ix = (int(rng.integers(p32.src_k + 3, k_exit + 1)), int(rng.integers(0, shape[1])),
              int(rng.integers(0, shape[2])))
        Vp = Vbase.clone()
        Vp[ix] += hstep
        Vm = Vbase.clone()
        Vm[ix] -= hstep
        with torch.no_grad():
            fd = (loss_of(Vp, p64).item() - loss_of(Vm, p64).item()) / (2 * hstep)
        if abs(fd) < resolvable:
            redrawn += 1
            continue
        rows.append((ix, fd, float(g64[ix]), float(g32[ix])))
        if len(rows) % 8 == 0:
            print(f"[adjoint] {len(rows)}/{n_fd} finite differences", file=sys.stderr,
                  flush=True)
    big = None if quick else _adjoint_timing((256, 512, 512), nsteps, samples)
    err64 = [abs(r[2] - r[1]) / abs(r[1]) for r in rows]
    err32 = [abs(r[3] - r[1]) / abs(r[1]) for r in rows]
    return {"study": "adjoint", "grid_nz_ny_nx": list(shape), "steps": nsteps,
            "loss": float(L0), "cost": costs,
            "backward_over_forward": costs["sqrt_t"]["backward_over_forward"],
            "fd_step": hstep, "fd_voxels": len(rows), "fd_redrawn": redrawn,
            "max_rel_err_f64_adjoint_vs_fd": float(max(err64)),
            "median_rel_err_f64_adjoint_vs_fd": float(np.median(err64)),
            "max_rel_err_f32_adjoint_vs_fd": float(max(err32)),
            "median_rel_err_f32_adjoint_vs_fd": float(np.median(err32)),
            "rel_l2_f32_vs_f64_gradient": rel_l2(g32, g64),
            "timing_512x512x256": big,
            "samples": [{"voxel": list(r[0]), "fd": r[1], "adj_f64": r[2], "adj_f32": r[3]}
                        for r in rows[:6]]}


def _time_forward_backward(loss_of, V, reps=3):
    """Best of reps wall times of loss_of(V) without gradient and of loss_of(V) plus its
    backward (the first backward also allocates the records); returns them and the last
    leaf tensor, whose .grad holds the gradient."""
    t_fwd = t_fb = math.inf
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            loss_of(V)
        torch.cuda.synchronize()
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    for _ in range(reps):
        Vg = V.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss_of(Vg).backward()
        torch.cuda.synchronize()
        t_fb = min(t_fb, time.perf_counter() - t0)
    return t_fwd, t_fb, Vg


def _adjoint_timing(shape, nsteps, samples, n_atoms=150):
    """Forward and forward + backward time of the adjoint study's loss on another grid
    (random target: timing only), with sqrt(T) checkpoints and with the default plan."""
    from .adjoint import exit_wave, far_field_intensity
    from .adjoint import last_plan as adjoint_plan
    p, V = kg_slab(shape, precision="f32", device="cuda", n_atoms=n_atoms, seed=2)
    p = p.with_(dt=1e-6)
    V32 = V.to(torch.float32)
    del V
    k_exit = p.nz - 16 - 3
    It = torch.rand((p.ny, p.nx), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))

    def loss_of(Vx, **kw):
        E = exit_wave(Vx, p, nsteps, k_exit, samples=samples, **kw)
        return torch.sum((far_field_intensity(E) - It) ** 2)

    with torch.no_grad():  # warm-up (allocations, library load)
        loss_of(V32)
    out = {"grid_nz_ny_nx": list(shape), "steps": nsteps}
    for plan_name, kw in (("sqrt_t", {"checkpoint_every": math.ceil(math.sqrt(nsteps))}),
                          ("memory_budget", {})):
        t_fwd, t_fb, _ = _time_forward_backward(lambda Vx: loss_of(Vx, **kw), V32)
        out[plan_name] = {"segment_plan": dict(adjoint_plan), "t_forward_s": t_fwd,
                          "t_forward_backward_s": t_fb,
                          "backward_over_forward": (t_fb - t_fwd) / t_fwd}
        print(f"[adjoint] {shape} {plan_name}: forward {t_fwd:.2f} s, forward+backward "
              f"{t_fb:.2f} s", file=sys.stderr, flush=True)
    return out


def _notebook_grid(dz_roi):
    """The plane-wave comparison grid of ``fdtd_3D_baseline_comparison.ipynb`` (cell 3):
    64 x 64 over [-1, 1]^2 A, the ROI z in [-1, 1] at spacing dz_roi, padded by a 0.4 A
    PML and a 0.2 A source buffer on each side."""
    nx = ny = 64
    roi_nz = int(round(2.0 / dz_roi)) + 1
    z = np.linspace(-1.0, 1.0, roi_nz)
    dz = z[1] - z[0]
    pml, buf = 0.4, 0.2
    pad = pml + buf
    ext_zmin = -1.0 - pad
    ext_nz = int(np.round((1.0 + pad - ext_zmin) / dz))
    z_ext = np.linspace(ext_zmin, ext_zmin + (ext_nz - 1) * dz, ext_nz)
    return {"nx": nx, "ny": ny, "roi_nz": roi_nz, "dz": dz, "z": z, "z_ext": z_ext,
            "pml": pml, "buf": buf, "pad": pad,
            "extent_roi": (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0),
            "extent_ext": (-1.0, 1.0, -1.0, 1.0, ext_zmin, float(z_ext[-1]))}


def fdtd_vs_multislice(dz_roi, steps, wavelength=0.05, Z=79.0, screening=0.3):
    """FDTD exit wave of a plane wave through one atom against the Fresnel and angular
    spectrum multislices of the same potential, following the notebook's recipe
    (``fdtd_3D_baseline_comparison.ipynb`` cells 3-16): the potential fills the ROI
    planes of the padded FDTD grid (zero elsewhere); the FDTD runs through the drop-in
    ``RelativisticSolver3D.run``; its exit plane (z = 1 A) is demodulated and normalised
    by the incident intensity two planes before the ROI; the multislices start from a
    plane wave pre-propagated (angular spectrum) over the padding."""
    from . import spectral as sp
    from .fdtd_solver import RelativisticSolver3D
    g = _notebook_grid(dz_roi)
    nx, ny, roi_nz, dz = g["nx"], g["ny"], g["roi_nz"], g["dz"]
    z_ext = g["z_ext"]
    ext_nz = len(z_ext)
    # the atom's potential energy on the ROI planes (eV, negative); the electrostatic
    # potential the multislice takes is its negative
    V_roi = atom_potential((roi_nz, ny, nx), g["extent_roi"], [(0.0, 0.0, 0.0)], [Z],
                           [screening], r_soft=2.0 / (nx - 1), device="cuda",
                           dtype=torch.float64)
    k0 = int(np.argmin(np.abs(z_ext - (-1.0))))
    k0 = min(k0, ext_nz - roi_nz)
    V_ext = np.zeros((ext_nz, ny, nx), dtype=np.float64)
    V_ext[k0:k0 + roi_nz] = V_roi.cpu().numpy()
    k_source = int((g["pml"] + 0.5 * g["buf"]) / dz)
    k_source = max(1, min(max(k_source, int(np.ceil(g["pml"] / dz)) + 2), ext_nz - 2))
    t0 = time.perf_counter()
    res = RelativisticSolver3D().run(V_ext, wavelength_angstrom=wavelength,
                                     extent=g["extent_ext"], n_frames=1, steps_per_frame=steps,
                                     snapshot_count=1, pml_thick=g["pml"], k_source=k_source,
                                     use_angstrom_units=True)
    t_fdtd = time.perf_counter() - t0
    if not res:
        raise RuntimeError("FDTD run failed")
    k_exit = int(np.argmin(np.abs(z_ext - 1.0)))
    k_mon = max(0, int(np.argmin(np.abs(z_ext - (-1.0)))) - 2)
    ex = np.zeros((ny, nx), dtype=np.complex128)
    mon = np.zeros((ny, nx), dtype=np.complex128)
    for t, phi, psi in res["snapshots"]:
        f = (phi.astype(np.float64) + 1j * psi.astype(np.float64)) * np.exp(1j * float(res["E_sim"]) * float(t))
        ex += f[k_exit]
        mon += f[k_mon]
    n = max(1, len(res["snapshots"]))
    ex, mon = ex / n, mon / n
    i_inc = float(np.mean(np.abs(mon) ** 2))
    if i_inc > 1e-9:
        ex = ex / math.sqrt(i_inc)
    # spectral side (energy of the beam from the wavelength with the spectral constants)
    pc = sp.HPLANCK * sp.C_LIGHT / (wavelength * 1e-10) / sp.E_CHARGE
    e0 = sp.electron_rest_energy()
    energy = math.sqrt(pc * pc + e0 * e0) - e0
    ps = (2.0 / (ny - 1), 2.0 / (nx - 1))
    probe = torch.ones((ny, nx), dtype=torch.complex128, device="cuda")
    probe = sp.Propagator(probe, sp.angular_spectrum_propagation_kernel(ny, nx, ps, g["pad"], energy))
    phi_roi = -V_roi
    out = {"grid_ext_nz_ny_nx": [ext_nz, ny, nx], "roi_planes": roi_nz, "dz": dz,
           "wavelength_A": wavelength, "energy_eV": energy, "fdtd_steps": steps,
           "k_source": k_source, "k_exit": k_exit, "fdtd_s": t_fdtd,
           "fdtd_gvoxel_steps_per_s": res.get("gvoxel_steps_per_s")}
    exits = {}
    for name, kern in (("fresnel", sp.fresnel_propagation_kernel),
                       ("angular_spectrum", sp.angular_spectrum_propagation_kernel)):
        H = kern(ny, nx, ps, dz, energy)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e_ms, _, _ = sp.simulate_fresnel_as(phi_roi, probe, H, dz, energy, keep_wavefronts=False)
        torch.cuda.synchronize()
        out[f"{name}_s"] = time.perf_counter() - t0
        exits[name] = e_ms
    fd = torch.tensor(ex, device="cuda")
    out["fdtd_vs_fresnel"] = sp.compare_exit_waves(fd, exits["fresnel"])
    out["fdtd_vs_angular_spectrum"] = sp.compare_exit_waves(fd, exits["angular_spectrum"])
    out["fresnel_vs_angular_spectrum"] = sp.compare_exit_waves(exits["fresnel"],
                                                               exits["angular_spectrum"])
    return out


def multislice_throughput(n=512, slices=256, dtype=torch.complex64, reps=3):
    """Slices per second of the GPU multislice (HIP phase grating + rocFFT) on an n x n
    plane wave through a random potential, and the phase-grating kernel's own rate."""
    from . import spectral as sp
    gen = torch.Generator(device="cuda").manual_seed(6)
    rdt = torch.float32 if dtype == torch.complex64 else torch.float64
    pot = 50.0 * torch.rand((slices, n, n), generator=gen, device="cuda", dtype=rdt)
    probe = torch.ones((n, n), dtype=dtype, device="cuda")
    e, dz, ps = 200e3, 0.05, (0.05, 0.05)
    H = sp.angular_spectrum_propagation_kernel(n, n, ps, dz, e)
    sp.simulate_fresnel_as(pot, probe, H, dz, e, keep_wavefronts=False)
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sp.simulate_fresnel_as(pot, probe, H, dz, e, keep_wavefronts=False)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    w = probe.clone()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for i in range(slices):
        sp.transmit(w, pot[i], dz, e)
    b.record()
    torch.cuda.synchronize()
    t_tr = a.elapsed_time(b) / 1e3 / slices
    nbytes = n * n * (2 * w.element_size() + pot.element_size())
    return {"n": n, "slices": slices, "dtype": str(dtype).replace("torch.", ""),
            "slices_per_s": slices / best, "ms_per_slice": best / slices * 1e3,
            "transmit_us": t_tr * 1e6, "transmit_GBs": nbytes / t_tr / 1e9}


def spectral_study(quick=False):
    """FDTD against spectral multislice exit waves, and the multislice's throughput."""
    rows = [fdtd_vs_multislice(2.0 / 127, 8000 if quick else 30000)]
    if not quick:
        rows.append(fdtd_vs_multislice(0.005, 30000))
    thr = [multislice_throughput(n=128 if quick else 512, slices=32 if quick else 256,
                                 dtype=d) for d in (torch.complex64, torch.complex128)]
    for r in rows:
        print(f"[spectral] dz {r['dz']:.4f}: fdtd vs AS {r['fdtd_vs_angular_spectrum']}",
              file=sys.stderr, flush=True)
    return {"study": "spectral", "comparisons": rows, "multislice_throughput": thr}


def api_study(quick=False):
    """Throughput of the drop-in entry points around the stepper: the reference's Python
    API (``RelativisticSolver3D.run`` and the 2D ``RelativisticSolver.run``, in process,
    one snapshot), the file-contract command (``backend="cli"``) against the in-process
    run on the same grid, and the GPU atom-potential generator."""
    import contextlib
    import io
    from .fdtd_solver import RelativisticSolver, RelativisticSolver3D
    from .workloads import atom_potential, centred_extent, random_slab
    out = {"study": "api"}
    quiet = contextlib.redirect_stdout(io.StringIO())   # the API prints like the reference

    # the atom-potential generator on the benchmark grid (400 atoms, one launch)
    shape = (128, 128, 256) if quick else (512, 256, 256)
    ext = centred_extent(shape, 0.05, 0.005)
    half = ext[5] - 20 * 0.005
    pos, Zs, scr = random_slab(400, ((ext[0], ext[1] + 0.05), (ext[2], ext[3] + 0.05)),
                               (-half, half), 0)
    best = math.inf
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        V = atom_potential(shape, ext, pos, Zs, scr, r_soft=0.05, device="cuda")
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    nvox = shape[0] * shape[1] * shape[2]
    out["atom_potential"] = {"grid_nz_ny_nx": list(shape), "atoms": 400, "ms": best * 1e3,
                             "Gvoxels_per_s": nvox / best / 1e9}

    # the reference's 3D API (Schrodinger, Angstrom / eV), host V in, one snapshot out
    Vh = V.cpu().numpy()
    nfr, spf = (4, 50) if quick else (10, 100)
    with quiet:
        t = time.perf_counter()
        r = RelativisticSolver3D().run(Vh, 0.025, ext, n_frames=nfr, steps_per_frame=spf,
                                       snapshot_count=1, pml_thick=0.08, k_source=18)
        wall = time.perf_counter() - t
    out["solver3d_run"] = {"grid_nz_ny_nx": list(shape), "steps": nfr * spf,
                           "stepping_Gvoxel_steps_per_s": r["gvoxel_steps_per_s"],
                           "wall_s": wall,
                           "wall_Gvoxel_steps_per_s": nvox * nfr * spf / wall / 1e9}

    # the 2D API (atomic units, y the propagation axis)
    ny2, nx2 = (1024, 512) if quick else (4096, 2048)
    V2 = np.zeros((ny2, nx2), dtype=np.float32)
    with quiet:
        t = time.perf_counter()
        r = RelativisticSolver().run(V2, 0.05, (-20.0, 20.0, -40.0, 40.0), n_frames=nfr,
                                     steps_per_frame=spf, snapshot_count=1)
        wall = time.perf_counter() - t
    out["solver2d_run"] = {"grid_ny_nx": [ny2, nx2], "steps": nfr * spf,
                           "stepping_Gvoxel_steps_per_s": r["gvoxel_steps_per_s"],
                           "wall_s": wall}

    # the file contract (potential.bin -> fdtd_3d_bin -> snapshots.bin) against the
    # in-process run of the same job
    cshape = (64, 64, 64)
    cext = centred_extent(cshape, 0.05)
    Vc = np.zeros(cshape, dtype=np.float32)
    times = {}
    for backend in ("hip", "cli"):
        with quiet:
            t = time.perf_counter()
            r = RelativisticSolver3D().run(Vc, 0.05, cext, n_frames=4, steps_per_frame=25,
                                           snapshot_count=4, pml_thick=0.4, k_source=10,
                                           backend=backend)
            times[backend] = time.perf_counter() - t
        assert r is not None and len(r["snapshots"]) == 4
    out["cli_vs_inprocess"] = {"grid_nz_ny_nx": list(cshape), "steps": 100, "snapshots": 4,
                               "inprocess_s": times["hip"], "cli_s": times["cli"]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["precision", "batched", "adjoint", "spectral", "api"])
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    fn = {"precision": precision_study, "batched": batched_study, "adjoint": adjoint_study,
          "spectral": spectral_study, "api": api_study}
    out = fn[a.study](quick=a.quick)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
