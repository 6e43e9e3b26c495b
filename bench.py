"""Throughput of the fused RK4 FDTD stepper on one MI355X.

A step is one RK4 step (soft source + the RK4 stages, fused into as few
passes as the schedule allows: one for Klein-Gordon float32) over the whole
grid.  Default workload: Klein-Gordon (psi, pi) plane wave through a
400-atom random slab on a 256 x 256 x 512 grid (dx = dy = 0.05 A, dz = 0.005 A,
PML 16 planes per z face), complex64 fields, float32 V.  Inputs are resident
in HBM before the timed region; the timed region is K steps bracketed by
``torch.cuda.synchronize()``.

Printed (one JSON line): Gvoxel-steps/s; the roofline of the dominant kernel
(algorithmic bytes per launch / its mean HIP-event duration, measured on the
solver's stream); the CPU baseline (the oracle's FP64 step restated in C +
OpenMP, on a bounded slab of the same workload, timed on this host's cores); the cost of the gradient's reverse
steps on the same grid (``adjoint``: a recorded forward step and an adjoint
step against a plain step, per step, and the backward / forward ratio of a
1500-step run checkpointed every ceil(sqrt(1500)) = 39 steps at those costs).

Roofline bytes.  ``achieved`` counts the algorithmic bytes of the schedule that
runs: what the dominant kernel must read and write per voxel (the fused
Klein-Gordon pass: read psi, pi, V, write psi', pi' = 36 B; float64 72 B; Kahan
adds the compensation's read and write), times the voxels of one launch, over
its mean duration; ``frac`` = achieved / 8 TB/s.  ``traffic`` is the HBM bytes per
launch that rocprofv3 measured for the kernel on this workload (FETCH_SIZE +
WRITE_SIZE, calibrated on a plain copy), read from the summary named in
``traffic_source``.  ``canonical_rate`` restates the step at the bytes of an RK4
step that streams once per stage (Klein-Gordon complex64: 272 B per voxel-step,
Schrodinger complex64: 144 B; FP64 doubles them; Kahan adds the compensation
buffer), the basis the throughput target is quoted on; a fused schedule moves
fewer bytes than that, so its ``x_peak`` may exceed 1.  ``measured_stream_*`` is
the best plain copy / read on the same device.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per voxel for each pass of a step
PASS_BYTES = {
    # one streaming pass per RK4 stage
    ("stages", "klein_gordon", "f32"): (52, 84, 84, 52),
    ("stages", "klein_gordon", "f32_kahan"): (52, 84, 84, 84),
    ("stages", "klein_gordon", "f64"): (104, 168, 168, 104),
    ("stages", "schrodinger", "f32"): (28, 44, 44, 28),
    ("stages", "schrodinger", "f32_kahan"): (28, 44, 44, 52),
    ("stages", "schrodinger", "f64"): (56, 88, 88, 56),
    # Klein-Gordon RK4 step in two stencil passes
    ("kg_2pass", "klein_gordon", "f32"): (60, 68),
    ("kg_2pass", "klein_gordon", "f32_kahan"): (60, 108),
    ("kg_2pass", "klein_gordon", "f64"): (120, 136),
    # both stencil sweeps in one pass
    ("kg_fused", "klein_gordon", "f32"): (36,),
    ("kg_fused", "klein_gordon", "f64"): (72,),
    ("kg_fused", "klein_gordon", "f32_kahan"): (68,),   # + the compensation, read and written
    # Schrodinger RK4 polynomial in Horner form: 2 passes of 2 stencil levels
    ("schr_fused", "schrodinger", "f32"): (20, 28),
    ("schr_fused", "schrodinger", "f64"): (40, 56),
    ("schr_fused", "schrodinger", "f32_kahan"): (20, 44),  # pass B + the compensation, R/W
    # the four Horner levels in one pass: read psi_n, V; write psi_{n+1} (f64: 2D grids)
    ("schr_1pass", "schrodinger", "f32"): (20,),
    ("schr_1pass", "schrodinger", "f64"): (40,),
    ("schr_1pass", "schrodinger", "f32_kahan"): (36,),   # + the compensation, read and written
}


# canonical bytes per voxel-step: one streaming pass per RK4 stage
CANONICAL_BYTES = {
    ("schrodinger", "f32"): 144, ("schrodinger", "f32_kahan"): 160, ("schrodinger", "f64"): 288,
    ("klein_gordon", "f32"): 272, ("klein_gordon", "f32_kahan"): 304, ("klein_gordon", "f64"): 544,
}


def schedule_of(p):
    return p.resolved_schedule()


def profiled_traffic(kernel, p):
    """(HBM bytes per launch of this kernel on this workload, the summary's path) from the
    committed rocprofv3 summary (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy);
    (None, None) if not profiled."""
    import re
    eq = "kg" if p.equation == "klein_gordon" else "schr"
    name = f"{kernel}_{eq}_{p.nx}x{p.ny}x{p.nz}_{p.precision}.md"
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch\s*=\s*(\d+)", f.read())
        return (int(m.group(1)), f"profiles/{name}") if m else (None, None)
    except OSError:
        return None, None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="f32", choices=["f32", "f32_kahan", "f64"])
    ap.add_argument("--equation", default="klein_gordon", choices=["klein_gordon", "schrodinger"])
    ap.add_argument("--shape", default="512,256,256", help="nz,ny,nx (ny = 1: a 2D grid)")
    ap.add_argument("--end-planes", default="frozen", choices=["frozen", "zero"],
                    help="Schrodinger end-plane rule: frozen (3D reference) or zero stage "
                         "inputs on the end rows (the 2D reference)")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--schedule", default="auto",
                    choices=["auto", "stages", "kg_2pass", "kg_fused", "schr_fused", "schr_1pass"])
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB (HBM counter calibration)")
    ap.add_argument("--adjoint-steps", type=int, default=16,
                    help="reverse steps timed for the adjoint figures (0: skip)")
    return ap.parse_args()


def adjoint_cost(sim, nsteps, slots=8):
    """Per-step times of the backward's two parts on this handle: a recorded forward step
    (the recompute of a checkpoint segment) and an adjoint step, with sqrt(T)-style slot
    reuse; the handle is advanced by the records."""
    import torch
    from wideanglepropagation_amd.adjoint import _Adjoint
    slots = max(1, min(slots, nsteps))
    adj = _Adjoint(sim, slots)
    try:
        for s in range(slots):       # warm-up
            adj.record(s, s)
            adj.step(slots - 1 - s)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for s in range(nsteps):
            adj.record(s, s % slots)
        torch.cuda.synchronize()
        rec = (time.perf_counter() - t) / nsteps
        t = time.perf_counter()
        for s in range(nsteps):
            adj.step((nsteps - 1 - s) % slots)
        torch.cuda.synchronize()
        rev = (time.perf_counter() - t) / nsteps
    finally:
        adj.close()
    return rec, rev


def copy_calibration(gib, reps=3):
    """A plain copy kernel of known size: reads and writes gib GiB per launch."""
    import torch
    n = int(gib * 2 ** 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t) / reps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "copy_GBs": 2 * n * 4 / el / 1e9}), flush=True)


def cpu_baseline(eq, shape, budget_s=10.0, max_steps=2000):
    """FP64 C + OpenMP restatement of the oracle's step (oracle/fdtd3d_omp.c, checked
    against the NumPy oracle in tests/test_oracle_c.py) on a bounded slab of the workload:
    the same lateral size, 1/8 of the planes; a warm-up step and two timed steps to size
    the sample, then as many steps as fit in about budget_s seconds in one call, on up to
    16 host threads."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import fdtd3d, fdtd3d_c
    from wideanglepropagation_amd.workloads import kg_slab, schrodinger_gaussian_64
    nz, ny, nx = shape
    sub = (max(8, nz // 8), ny, nx)
    if eq == "klein_gordon":
        p, V = kg_slab(sub, precision="f64", n_atoms=50, device="cpu")
        V = V.numpy()
        st = fdtd3d.zero_state(p.oracle_dict())
    else:
        p, V, st, _ = schrodinger_gaussian_64("f64")
        sub = p.shape
    d = p.oracle_dict()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = max(1, min(16, cores))
    st = fdtd3d_c.run(st, V, 1, d, threads=threads)       # warm-up (library, pages)
    t = time.perf_counter()
    st = fdtd3d_c.run(st, V, 2, d, step0=1, threads=threads)
    one = (time.perf_counter() - t) / 2
    nsteps = max(1, min(max_steps, int(budget_s / max(one, 1e-6))))
    t = time.perf_counter()
    st = fdtd3d_c.run(st, V, nsteps, d, step0=3, threads=threads)
    el = time.perf_counter() - t
    vox = sub[0] * sub[1] * sub[2]
    return {"value": vox * nsteps / el / 1e9, "unit": "Gvoxel-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"FP64 C + OpenMP restatement of the oracle's RK4 step "
                      f"(oracle/fdtd3d_omp.c), {eq} {sub[0]}x{sub[1]}x{sub[2]} (nz,ny,nx), "
                      f"{nsteps} step(s) in {el:.1f} s, {threads} threads of {cores} host cores"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU "
              "sharding is not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    import torch
    from wideanglepropagation_amd.engine import FDTD3D
    from wideanglepropagation_amd.workloads import kg_slab, centred_extent
    from wideanglepropagation_amd.physics import schrodinger_params, wavelength_from_kinetic_energy

    shape = tuple(int(v) for v in args.shape.split(","))
    nz, ny, nx = shape
    print(f"[bench] {args.equation} {args.precision} grid {nx}x{ny}x{nz}", file=sys.stderr,
          flush=True)
    if args.equation == "klein_gordon":
        p, V = kg_slab(shape, precision=args.precision, device="cuda")
    else:
        lam = wavelength_from_kinetic_energy(200e3)
        ext = centred_extent(shape, 0.05)
        p = schrodinger_params(lam, ext, shape, pml_thick=16 * 0.05, k_source=18,
                               precision=args.precision)
        from wideanglepropagation_amd.workloads import atom_potential, random_slab
        pos, Zs, scr = random_slab(400, ((ext[0], ext[1]), (ext[2], ext[3])), (-10, 10), 0)
        if ny == 1:
            # a 2D grid: the atoms in the x-z plane, as row 0 of a two-row grid whose other
            # row (and the periodic images) lie 100 A away, beyond the cutoff
            pos[:, 1] = 0.0
            ext2 = (ext[0], ext[1], 0.0, 100.0, ext[4], ext[5])
            V = atom_potential((nz, 2, nx), ext2, pos, Zs, scr, 0.05, device="cuda")
            V = V[:, :1, :].contiguous()
        else:
            V = atom_potential(shape, ext, pos, Zs, scr, 0.05, device="cuda")
    p = p.with_(tile_y=args.tile_y, tile_z=args.tile_z, schedule=args.schedule)
    if args.end_planes != "frozen":
        if args.equation != "schrodinger":
            sys.exit("--end-planes zero is the Schrodinger 2D rule")
        p = p.with_(end_planes=args.end_planes)
    V = V.to(torch.float64 if args.precision == "f64" else torch.float32).contiguous()
    sim = FDTD3D(p, V=V)
    del V
    ty, tz, groups = sim.launch_shape()
    print(f"[bench] launch tile_y={ty} tile_z={tz} workgroups/pass={groups} "
          f"device bytes={sim.device_bytes() / 2**30:.2f} GiB", file=sys.stderr, flush=True)

    sim.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    vox = p.voxels
    value = vox * args.steps / elapsed / 1e9
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-pass HIP-event timing on the solver's stream
    pass_ms, tot_ms = sim.step_profiled(args.profile_steps)
    n_prof = max(1, args.profile_steps)
    sched = schedule_of(p)
    pb = PASS_BYTES[(sched, p.equation, p.precision)]
    avg = [m / n_prof for m in pass_ms][:len(pb)]
    dom = int(np.argmax(avg))  # longest pass
    comp_bytes = pb[dom] * vox
    canon_bytes = int(round(CANONICAL_BYTES[(p.equation, p.precision)] * vox * pb[dom] / sum(pb)))
    dom_s = avg[dom] * 1e-3
    canon_gbs = canon_bytes / dom_s / 1e9
    comp_gbs = comp_bytes / dom_s / 1e9
    step_bytes = sum(pb) * vox
    step_gbs = step_bytes / (sum(avg) * 1e-3) / 1e9
    norm = sim.norm2()
    adjoint = None
    if args.adjoint_steps > 0 and p.precision != "f32_kahan" and p.end_planes == "frozen":
        rec_s, rev_s = adjoint_cost(sim, args.adjoint_steps)
        step_s = elapsed / args.steps
        T = 1500
        S = int(np.ceil(np.sqrt(T)))
        adjoint = {"steps": args.adjoint_steps, "step_ms": round(step_s * 1e3, 4),
                   "record_ms": round(rec_s * 1e3, 4), "adjoint_ms": round(rev_s * 1e3, 4),
                   "backward_over_forward": round((rec_s + rev_s) / step_s, 3),
                   # T steps checkpointed every S: the backward recomputes the T - S steps
                   # of every segment but the last (recorded in the forward) and reverses T
                   "sqrt_t_plan": {"T": T, "S": S, "recomputed_steps": T - S,
                                   "backward_over_forward": round(
                                       ((T - S) * rec_s + T * rev_s) / (T * step_s), 3)},
                   "adjoint_kernel": ("schr_adj1_pass (1 pass)" if sched == "schr_1pass" else
                                      "schr_adj_pass (2 passes)" if sched == "schr_fused" else
                                      "kg1_adj_pass" if p.equation == "klein_gordon"
                                      and vox * 8 < 2 ** 31 else "adj_pass (4 stage passes)"),
                   "adjoint_Gvoxels": round(vox / rev_s / 1e9, 2)}
        print(f"[bench] adjoint: record {rec_s * 1e3:.3f} ms, reverse step {rev_s * 1e3:.3f} ms "
              f"per step", file=sys.stderr, flush=True)
    sim.close()
    from wideanglepropagation_amd._lib import stream_bandwidth
    copy_gbs, read_gbs = stream_bandwidth(2 << 30, 10)

    one_pass = "schr2d_pass" if ny == 1 else "schr1_pass"
    kernel = {"kg_2pass": f"kg2_pass_p{dom + 1}", "kg_fused": "kg1_pass", "schr_1pass": one_pass,
              "schr_fused": f"schr2_pass_p{dom + 1}"}.get(sched, f"rk4_pass_s{dom + 1}")
    # the Schrodinger summary covers both passes and quotes pass B's bytes
    traffic, traffic_src = (profiled_traffic("schr2_pass", p) if sched == "schr_fused" and dom == 1
                            else profiled_traffic(kernel, p))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p.equation, shape)

    out = {
        "metric": "Gvoxel-steps/s",
        "value": round(value, 4),
        "unit": "Gvoxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if p.precision == "f64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{p.equation} RK4 FDTD {nx}x{ny}x{nz}, "
                               f"{'complex128' if p.precision == 'f64' else 'complex64'} fields "
                               f"({p.precision}), 400-atom slab, PML 16 planes, 200 keV"
                               + (", zero stage inputs on the end rows (the 2D rule)"
                                  if p.end_planes == "zero" else ""),
                   "grid": [nz, ny, nx], "precision": p.precision,
                   "equation": p.equation, "schedule": sched, "tile_y": ty, "tile_z": tz},
        "roofline": {"bound": "hbm", "achieved": round(comp_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(comp_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": {"kg_2pass": f"kg2_pass pass {dom + 1}",
                                "kg_fused": "kg1_pass", "schr_1pass": one_pass,
                                "schr_fused": f"schr2_pass pass {'AB'[dom]}"}.get(
                                    sched, f"rk4_pass stage {dom + 1}"),
                     "basis": f"algorithmic {pb[dom]} B per voxel of the executed {sched} "
                              f"schedule's dominant kernel x {vox} voxels per launch",
                     "bytes_per_launch": comp_bytes,
                     "mean_launch_ms": round(avg[dom], 5),
                     "time_source": "HIP events on the solver's stream",
                     "traffic_source": (f"{traffic_src} (rocprofv3 FETCH_SIZE + WRITE_SIZE of "
                                        "this kernel and workload, copy-calibrated)"
                                        if traffic else None),
                     "traffic_achieved": (round(traffic / dom_s / 1e9, 1) if traffic else None),
                     "traffic_frac": (round(traffic / dom_s / 1e9 / HBM_PEAK_GBS, 4)
                                      if traffic else None)},
        "canonical_rate": {"basis": f"{CANONICAL_BYTES[(p.equation, p.precision)]} B per "
                                    "voxel-step (one streaming pass per RK4 stage), apportioned "
                                    "to the dominant kernel by its share of the step's bytes",
                           "bytes_per_launch": canon_bytes, "GBs": round(canon_gbs, 1),
                           "x_peak": round(canon_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "detail": {"pass_ms": [round(a, 5) for a in avg],
                   "algorithmic_bytes_per_voxel_step": sum(pb),
                   "step_effective_GBs": round(step_gbs, 1),
                   "measured_stream_copy_GBs": round(copy_gbs, 1),
                   "measured_stream_read_GBs": round(read_gbs, 1),
                   "norm2_final": norm},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
