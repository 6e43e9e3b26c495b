"""The batched workload's grids against the oracle.

One instance of each size class of the batched-probe workload (96^3, 128^3,
160 x 160 x 256, 192^3; dx = dy = 0.05 A, dz = 0.005 A), with mixed equations
(Schrodinger on 96^3 and 160 x 160 x 256, Klein-Gordon on 128^3 and 192^3), beam
energies drawn from U[60, 300] keV, a convergent Gaussian probe at a random
lateral position injected through the soft source, and the shared 200-atom slab.
The batch is stepped through its captured hipGraphs (complex64, the fused
schedules), and each instance is compared with the FP64 C + OpenMP oracle
(oracle/fdtd3d_omp.c, pinned to the NumPy oracle) on the same inputs: float32
tolerance 3e-6 relative L2 per state component.  A random initial state makes
every voxel of every grid carry signal from the first step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CLASSES = [((96, 96, 96), "schrodinger"), ((128, 128, 128), "klein_gordon"),
           ((256, 160, 160), "schrodinger"), ((192, 192, 192), "klein_gordon")]


def test_batched_size_classes_against_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import fdtd3d_c
    from wideanglepropagation_amd.batched import BatchedFDTD
    from wideanglepropagation_amd.engine import FDTD3D
    from wideanglepropagation_amd.physics import (klein_gordon_params, schrodinger_params,
                                                  wavelength_from_kinetic_energy)
    from wideanglepropagation_amd.studies import probe_profile
    from wideanglepropagation_amd.workloads import atom_potential, centred_extent, random_slab
    rng = np.random.default_rng(1)
    spacing, spacing_z = 0.05, 0.005
    big = max(c[0][1] for c in CLASSES) * spacing
    pos, Zs, scr = random_slab(200, ((-big / 2, big / 2), (-big / 2, big / 2)), (-0.2, 0.2), 11)
    cases = []
    for shape, eq in CLASSES:
        ekin = float(rng.uniform(60e3, 300e3))
        lam = wavelength_from_kinetic_energy(ekin)
        ext = centred_extent(shape, spacing, spacing_z)
        mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
        p = mk(lam, ext, shape, pml_thick=16 * spacing_z, k_source=18, precision="f32")
        V = atom_potential(shape, ext, pos, Zs, scr, r_soft=spacing, cutoff=4.0, device="cuda",
                           dtype=torch.float32).cpu().numpy()
        prof = probe_profile(shape, ext, rng.uniform(ext[0], ext[1]),
                             rng.uniform(ext[2], ext[3]), lam).astype(np.complex64)
        psi0 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        pi0 = (-1j * p.e_sim * psi0).astype(np.complex64) if eq == "klein_gordon" else None
        cases.append((p, V, prof, psi0, pi0))
    sims = [FDTD3D(p, V=V, psi0=psi0, pi0=pi0, source_profile=prof)
            for p, V, prof, psi0, pi0 in cases]
    assert {s.params.resolved_schedule() for s in sims} == {"schr_1pass", "kg_fused"}
    nsteps = 10
    try:
        b = BatchedFDTD(sims)
        print(f"case batched workload classes, {len(b.groups)} launch groups", flush=True)
        b.capture(5)
        b.replay()
        b.replay()
        b.synchronize()
        for s, (p, V, prof, psi0, pi0) in zip(sims, cases):
            got = s.get_state(np.complex128)
            got = got if isinstance(got, tuple) else (got,)
            st0 = psi0.astype(np.complex128)
            if pi0 is not None:
                st0 = (st0, pi0.astype(np.complex128))
            d = p.with_(precision="f64").oracle_dict(src_profile=prof.astype(np.complex128))
            want = fdtd3d_c.run(st0, V.astype(np.float64), nsteps, d, threads=16)
            want = want if isinstance(want, tuple) else (want,)
            rel = max(float(np.linalg.norm(g - w) / np.linalg.norm(w)) for g, w in zip(got, want))
            print(f"{p.equation} {p.shape}: rel L2 vs oracle {rel:.3e}", flush=True)
            assert rel < 3e-6, (p.shape, rel)
        b.close()
    finally:
        for s in sims:
            s.close()
