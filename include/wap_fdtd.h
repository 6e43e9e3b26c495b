/*
 * wap_fdtd.h -- C ABI of the MI355X (gfx950) 3D RK4 FDTD electron-wave solver.
 *
 * The reference has no library interface for this path: its Python driver
 * (wide_angle_propagation/fdtd_solver.py) rewrites #defines into a CUDA
 * source, compiles it with nvcc, writes potential.bin, runs the program with
 * argv [n_frames, steps_per_frame, snapshot_count, -|lambda|] and parses
 * snapshots.bin.  Each entry point below names the part of that contract it
 * replaces (file:line in the reference).  Everything is in-process, plain
 * pointers and sizes; no torch types cross this boundary.
 *
 * Conventions
 *   - Arrays are C order (nz, ny, nx), x contiguous (fdtd_3d.cu:81-83).
 *   - Complex data is interleaved (re, im): c64 = 2 x float, c128 = 2 x double.
 *     The reference's split phi/psi float arrays are phi = Re, psi = Im.
 *   - Every call returns WAP_OK (0) or a negative status; the message of the
 *     last failure on the calling thread is wap_last_error().  Nothing calls
 *     exit() or abort() (the reference's gpuAssert exits, fdtd_3d.cu:71-78).
 *   - Device work is ordered on the handle's stream (wap_set_stream, default
 *     the null stream).  Calls that return data to host memory synchronise it.
 *   - A handle is not thread-safe; distinct handles are independent.
 */
#ifndef WAP_FDTD_H
#define WAP_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WAP_ABI_VERSION 1

/* status codes */
#define WAP_OK 0
#define WAP_ERR_ARG (-1)
#define WAP_ERR_HIP (-2)
#define WAP_ERR_STATE (-3)
#define WAP_ERR_OOM (-4)

/* equations */
#define WAP_EQ_SCHRODINGER 0  /* i dpsi/dt = (-A Lap + V) psi   (fdtd_3d.cu:242-248) */
#define WAP_EQ_KLEIN_GORDON 1 /* first-order (psi, pi) system; no reference counterpart */

/* arithmetic of the stepper */
#define WAP_PREC_F64 0        /* complex128 fields, float64 V */
#define WAP_PREC_F32 1        /* complex64 fields, float32 V (reference precision) */
#define WAP_PREC_F32_KAHAN 2  /* complex64 fields + compensated state update */

/* element types of caller buffers */
#define WAP_DT_F32 0
#define WAP_DT_F64 1
#define WAP_DT_C64 2
#define WAP_DT_C128 3

/* time of step n */
#define WAP_TIME_EXACT 0  /* t = t0 + (step0 + n) dt in float64 */
#define WAP_TIME_F32 1    /* t = float((float)(step0 + n) * (float)dt), as fdtd_3d.cu:365;
                             t0 must be 0 (WAP_ERR_ARG otherwise) */

/* pass schedule of a step */
#define WAP_SCHED_AUTO 0      /* KG: KG_FUSED (every precision) while a plane is under
                               * 2^31 / 10 bytes (a chunk window of the fused passes, at least
                               * 10 planes, under 2 GiB), else KG_2PASS; Schrodinger with such
                               * planes:
                               * SCHR_1PASS (f32_kahan on 3D grids only), else
                               * SCHR_FUSED with frozen end planes; else STAGES */
#define WAP_SCHED_STAGES 1    /* one streaming pass per RK4 stage (4 per step) */
#define WAP_SCHED_KG_2PASS 2  /* the same RK4 polynomial in 2 stencil passes */
#define WAP_SCHED_KG_FUSED 3  /* both stencil passes in one pass (LDS halo, ping-pong state;
                               * f32_kahan updates its compensation in place) */
#define WAP_SCHED_SCHR_FUSED 4  /* Schrodinger, frozen end planes, every precision: the RK4
                                 * polynomial in Horner form, two passes of two stencil levels
                                 * each */
#define WAP_SCHED_SCHR_1PASS 5  /* Schrodinger f32 and f64 (f32_kahan on 3D grids, updating
                                 * its compensation in place), frozen or zero end planes: the
                                 * four Horner levels in one pass (20 B per voxel-step,
                                 * f32_kahan 36, f64 40; 2D grids: one wave per 56-column
                                 * tile and z chunk) */

/* planes k = 0 and nz-1 (never stepped) */
#define WAP_ENDS_FROZEN 0  /* stage inputs there = the injected state (fdtd_3d.cu:184-190) */
#define WAP_ENDS_ZERO 1    /* stage inputs there = 0 (2D reference, fdtd.cu:189-198);
                              Schrodinger, 4-pass schedule only */

#define WAP_MAX_SRC_BAND 8

/*
 * Run description.  Replaces the compile-time #defines of fdtd_3d.cu:17-67
 * that fdtd_solver.py:285-374 rewrites, and the DT macro (fdtd_3d.cu:30).
 */
typedef struct wap_desc {
  int32_t nx, ny, nz;         /* NX NY NZ */
  int32_t equation;           /* WAP_EQ_* */
  int32_t precision;          /* WAP_PREC_* */
  int32_t time_mode;          /* WAP_TIME_* */
  int32_t src_enabled;        /* 0: no soft source (the reference always injects) */
  int32_t src_k;              /* K_SOURCE */
  int32_t src_band;           /* SRC_BAND (<= WAP_MAX_SRC_BAND) */
  int32_t tile_y;             /* rows per workgroup (64 x tile_y threads); 0 = auto.
                                 KG_FUSED uses 4, 8 or 16 rows (auto 16) */
  int32_t tile_z;             /* planes per thread column; 0 = auto */
  int32_t schedule;           /* WAP_SCHED_*: how a step is split into passes */
  int32_t end_planes;         /* WAP_ENDS_*: planes k = 0, nz-1 */
  int32_t reserved0;
  double xmin, xmax, ymin, ymax, zmin, zmax; /* XMIN..ZMAX; spacing = extent/(n-1);
                                               ny = 1 is a 2D grid (no y term) */
  double dt;                  /* DT */
  double lap_coeff;           /* LAP_COEFF = A (Schrodinger) */
  double e_sim;               /* carrier frequency of the source phase (E_SIM) */
  double k_beam;              /* source wavenumber sqrt(E_sim/A) (fdtd_3d.cu:139) */
  double hbar_c;              /* KG: hbar c [eV A] */
  double rest_energy;         /* KG: m c^2 [eV] */
  double pml_thick, pml_sigma0, pml_theta, pml_power; /* PML_* (fdtd_3d.cu:64-67) */
  double kg_gamma_max;        /* KG sponge damping rate at the outer face */
  double src_alpha;           /* SRC_ALPHA */
} wap_desc;

typedef struct wap_handle wap_handle;

/* Library ABI version (WAP_ABI_VERSION). */
int wap_abi_version(void);

/* Thread-local message of the last failure (replaces fdtd_3d.cu:71-78 stderr + exit). */
const char *wap_last_error(void);

/* Allocate device state for one grid (replaces fdtd_3d.cu:277-344: host/device
 * allocation, zero initial field, launch geometry).  The field starts at 0. */
int wap_create(const wap_desc *desc, wap_handle **out);

/* Release every device buffer of the handle (fdtd_3d.cu:400 frees only 3 of 7). */
int wap_destroy(wap_handle *h);

/* Order all later work of the handle on this hipStream_t (NULL = null stream). */
int wap_set_stream(wap_handle *h, void *stream);

/* Upload V(r) (nz,ny,nx), real, eV (replaces potential.bin: fdtd_solver.py:416-419,
 * fdtd_3d.cu:284-296).  dtype WAP_DT_F32 or WAP_DT_F64; on_device != 0 means the
 * pointer is device memory.  Must be called before wap_step. */
int wap_set_potential(wap_handle *h, const void *v, int dtype, int on_device);

/* on_device of wap_set_state / wap_get_state: 0 host memory, 1 device memory (the call
 * returns when the copy is done), WAP_STREAM_ORDERED device memory copied on the
 * handle's stream without waiting: the caller keeps the memory alive and orders its own
 * accesses after the handle's stream (checkpoints that only this handle reads back). */
#define WAP_STREAM_ORDERED 2

/* Set the state (replaces the zero init of fdtd_3d.cu:322-326).  psi (and pi for
 * Klein-Gordon; NULL means 0) are complex (nz,ny,nx) of dtype WAP_DT_C64 or C128.
 * psi == NULL resets the whole state to zero. */
int wap_set_state(wap_handle *h, const void *psi, const void *pi, int dtype, int on_device);

/* Copy the state out (replaces the per-frame D2H of fdtd_3d.cu:386-388).  pi may be
 * NULL.  dtype WAP_DT_C64 or C128; the F32_KAHAN state is exported as hi - comp. */
int wap_get_state(wap_handle *h, void *psi, void *pi, int dtype, int on_device);

/* Lateral profile P(x,y) of the soft source, complex (ny,nx); NULL = plane wave. */
int wap_set_source_profile(wap_handle *h, const void *prof, int dtype, int on_device);

/* Advance nsteps steps (replaces the hot loop fdtd_3d.cu:363-382: memsets,
 * inject_source_3d, 4 x rk4_stage_3d, rk4_final_3d).  Step n injects at
 * t_n (time_mode) and then takes one classical RK4 step.  Asynchronous. */
int wap_step(wap_handle *h, int64_t nsteps, int64_t step0, double t0);

/* As wap_step, timing every pass with HIP events on the handle's stream.
 * pass_ms[4] receives the summed device time of passes 1..4, total_ms the whole
 * region; synchronises. */
int wap_step_profiled(wap_handle *h, int64_t nsteps, int64_t step0, double t0,
                      double *pass_ms, double *total_ms);

/* Copy plane k of psi, complex (ny,nx), in dtype (C64/C128). */
int wap_get_plane(wap_handle *h, int k, void *out, int dtype, int on_device);

/* acc[j,i] += psi[k,j,i] * (cos_p + i sin_p) in float64 on device; acc is a device
 * complex128 (ny,nx) buffer.  The demodulated running average of the notebooks'
 * exit-wave recipe (fdtd_3D_baseline_comparison.ipynb:267-290) without dumping volumes. */
int wap_accumulate_plane(wap_handle *h, int k, double cos_p, double sin_p, void *acc);

/* sum |psi|^2 over the grid, accumulated in float64 on device; synchronises. */
int wap_norm2(wap_handle *h, double *out);

/* Wait for the handle's stream. */
int wap_synchronize(wap_handle *h);

/* Device bytes held by the handle. */
int wap_device_bytes(wap_handle *h, int64_t *out);

/* Effective launch shape chosen for the handle (tile_y, tile_z, workgroups per pass). */
int wap_launch_shape(wap_handle *h, int32_t *tile_y, int32_t *tile_z, int64_t *groups);

/*
 * Batches: many independent grids (the reference runs one grid per process)
 * stepped by one launch per pass.  All handles must share equation, precision
 * and schedule and have V set; sizes, energies, dt and sources may differ.
 * Fused schedules (KG_FUSED, SCHR_FUSED) run their fused passes over all grids;
 * they ping-pong each grid's state, so after an odd number of batched steps
 * (or replays of an odd-length graph) the handles' state buffers are swapped.
 * The batch keeps pointers to the handles' buffers: destroy it before them.
 */
typedef struct wap_batch wap_batch;

/* Group n handles; tile_y = workgroup rows for all (0 = the first handle's; the
 * fused schedules take 4, 8 or 16 waves, the others 1, 2, 4 or 8 rows). */
int wap_batch_create(wap_handle *const *handles, int n, int tile_y, wap_batch **out);
int wap_batch_destroy(wap_batch *b);
/* Stream of the batch's launches (default: the first handle's stream). */
int wap_batch_set_stream(wap_batch *b, void *stream);
/* Total workgroups per pass and the shared tile_y (fused: waves per workgroup). */
int wap_batch_tiles(wap_batch *b, int64_t *tiles, int32_t *tile_y);
/* Advance every grid nsteps steps (step n of grid i injects at its own t_n). */
int wap_batch_step(wap_batch *b, int64_t nsteps, int64_t step0);
/* Record nsteps steps of every grid into one hipGraph (on a non-null stream;
 * one is created if the batch has none). */
int wap_batch_capture(wap_batch *b, int64_t nsteps);
/* Run the captured graph for steps step0 .. step0+nsteps-1 (the per-step source
 * phasors are rewritten before the launch); ms != NULL returns its device time
 * (HIP events on the batch stream) and synchronises. */
int wap_batch_replay(wap_batch *b, int64_t step0, double *ms);
int wap_batch_synchronize(wap_batch *b);

/*
 * Batch sets: one hipGraph over several batches, e.g. one per equation of a mixed
 * workload (a batch holds one equation, precision and schedule).  Inside the graph each
 * batch's steps run on a stream of their own, forked from and joined into the set's
 * launch stream, so independent groups overlap on the device.  The set keeps its own
 * per-step phase tables; the batches and their handles must outlive it, and no other
 * work may step the same handles while a replay runs.
 */
typedef struct wap_batch_set wap_batch_set;
int wap_batch_set_create(wap_batch *const *batches, int n, wap_batch_set **out);
int wap_batch_set_destroy(wap_batch_set *s);
/* Record nsteps steps of every batch into the set's graph. */
int wap_batch_set_capture(wap_batch_set *s, int64_t nsteps);
/* Run the graph for steps step0 .. step0+nsteps-1 of every batch; ms != NULL returns its
 * device time (HIP events on the launch stream) and synchronises. */
int wap_batch_set_replay(wap_batch_set *s, int64_t step0, double *ms);
int wap_batch_set_synchronize(wap_batch_set *s);

/*
 * Discrete adjoint (no reference counterpart: the reference has no gradient).
 * Gradient of a real loss of exit-plane values with respect to V, through the
 * RK4 steps, the PML / sponge and the soft source.  Use: run the forward with
 * checkpoints; then per segment restore a checkpoint with wap_set_state,
 * wap_adjoint_record each step of the segment into slots 0..m-1 (this also
 * advances the handle), and walk the segment backwards with wap_adjoint_step
 * (slot m-1 .. 0), adding exit-plane seeds with wap_adjoint_seed.  f64 and f32
 * handles with frozen end planes.  Klein-Gordon grids within the fused pass's field
 * limit and Schrodinger handles on the fused schedule record each step's input state
 * (the record step is the fused forward step, whose output buffer is swapped with the
 * slot) and reverse a step in one (KG) or two (Schrodinger) fused passes; other
 * handles record the 4 stage inputs and reverse stage by stage.  Destroy the adjoint
 * before its handle.
 */
typedef struct wap_adjoint wap_adjoint;

/* Adjoint workspace with `slots` per-step records (fused schedules: the input state;
 * otherwise the 4 stage inputs). */
int wap_adjoint_create(wap_handle *h, int slots, wap_adjoint **out);
/* Device bytes wap_adjoint_create_in needs for `slots` records on this handle. */
int wap_adjoint_workspace_bytes(wap_handle *h, int slots, int64_t *bytes);
/* As wap_adjoint_create, with every device buffer carved from the caller's `workspace`
 * (device memory, 256-byte aligned, at least wap_adjoint_workspace_bytes): for callers
 * with a caching allocator, so that repeated backward passes neither allocate nor free
 * gigabytes of records.  The workspace must outlive the adjoint. */
int wap_adjoint_create_in(wap_handle *h, int slots, void *workspace, int64_t bytes,
                          wap_adjoint **out);
/* Gives the handle back its own buffers (copying the state if a record swap left it in
 * a record buffer), synchronises the handle's stream and releases the adjoint. */
int wap_adjoint_destroy(wap_adjoint *a);
/* lambda = 0, dL/dV = 0. */
int wap_adjoint_reset(wap_adjoint *a);
/* One forward step (its time as wap_step(h, 1, step, t0)), recording what its
 * reverse step needs into `slot`. */
int wap_adjoint_record(wap_adjoint *a, int64_t step, double t0, int slot);
/* lambda <- adjoint of the recorded step; dL/dV += its contribution. */
int wap_adjoint_step(wap_adjoint *a, int slot);
/* lambda_psi[k] += conj(c) g for a device complex128 (ny,nx) plane g. */
int wap_adjoint_seed(wap_adjoint *a, int k, const void *g, double c_re, double c_im);
/* dL/dV as float64 (nz,ny,nx); synchronises. */
int wap_adjoint_get_grad(wap_adjoint *a, void *out, int on_device);

/* Potential of screened-Coulomb atoms on a grid (replaces the abTEM potentials the
 * reference notebooks build, FDTDmethod/fdtd_3D_baseline_comparison.ipynb:129-180):
 * V = sum over atoms and their periodic x, y images within cutoff*s of
 * -14.399645 Z exp(-r/s) / sqrt(r^2 + r_soft^2) eV.  atoms: host float64 records
 * (x, y, z, Z, s) in A; extent (xmin, xmax, ymin, ymax, zmin, zmax); out: device
 * (nz, ny, nx) of dtype WAP_DT_F32 / F64, written on `stream`; synchronises. */
int wap_atom_potential(const double *atoms, int n_atoms, int nx, int ny, int nz,
                       const double *extent, double r_soft, double cutoff, void *out, int dtype,
                       void *stream);

/* As wap_atom_potential, allocating nothing: the atom records are staged in a static
 * device array of the library (WAP_STATIC_MAX_ATOMS records, 160 KiB per GPU, held until
 * the process ends) instead of a per-call device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code that a caller captures
 * into a HIP graph, tight serving loops, harnesses that forbid hipMalloc).  n_atoms above
 * WAP_STATIC_MAX_ATOMS is rejected (WAP_ERR_ARG).  Every call on one GPU shares that GPU's
 * staging array: calls that overlap on one GPU (other threads, other streams) race;
 * calls from one thread are ordered, since each synchronises `stream` before it returns. */
#define WAP_STATIC_MAX_ATOMS 4096
int wap_atom_potential_static(const double *atoms, int n_atoms, int nx, int ny, int nz,
                              const double *extent, double r_soft, double cutoff, void *out,
                              int dtype, void *stream);

/* Phase grating of one potential slice of a spectral multislice: psi *= exp(i 2 pi
 * (n(v) - 1) dz / lambda), n the relativistic electron refractive index at the
 * electrostatic potential v [V] for kinetic energy energy_ev (rest energy
 * rest_energy_ev).  Replaces electron_refractive_index and the phase grating of
 * simulate_fresnel_as (propagation.py:66-81, 330-335).  psi: n complex of dtype
 * WAP_DT_C64 / C128, v: n reals of dtype WAP_DT_F32 / F64, both device memory;
 * asynchronous on `stream` (NULL: the null stream). */
int wap_slice_transmit(void *psi, int psi_dtype, const void *v, int v_dtype, int64_t n,
                       double energy_ev, double rest_energy_ev, double slice_thickness,
                       double wavelength, void *stream);

/* Measured HBM bandwidth reference (no reference counterpart): a STREAM copy
 * (read + write) and a read-only sweep of `bytes`-sized device buffers, reps
 * launches each, timed with HIP events on the null stream; GB/s = 1e9 B/s. */
int wap_stream_bandwidth(int64_t bytes, int reps, double *copy_gbs, double *read_gbs);

#ifdef __cplusplus
}
#endif

#endif /* WAP_FDTD_H */
