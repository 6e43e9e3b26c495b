"""The HIP path against the oracle at the benchmark's full grid.

256 x 256 x 512 (nx, ny, nz; 33.5 M voxels): the grid of bench.py's default run, whose
fields outgrow the Infinity Cache, so the fused passes take their non-temporal-store
copies and run 1280 workgroups over 16 z chunks.  A random state goes through 3 steps of
each default schedule (Klein-Gordon fused in float32, float32 + Kahan and float64;
Schrodinger fused in float32, stage passes in float64).  The result is compared with
the FP64 C + OpenMP restatement of the oracle's step (oracle/fdtd3d_omp.c, itself pinned
to the NumPy oracle in tests/test_oracle_c.py) on the same inputs.  Tolerances:
float64 1e-12, float32 3e-6 relative L2 (as for the small grids).  The Klein-Gordon
float32 pass is also checked on 512 x 512 x 256 (2 steps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (512, 256, 256)
NSTEPS = 3


def _params(eq, precision):
    from wideanglepropagation_amd.physics import klein_gordon_params, schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent
    mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
    return mk(0.05, centred_extent(SHAPE, 0.05, 0.005), SHAPE, pml_thick=0.08, k_source=18,
              precision=precision)


@pytest.fixture(scope="module")
def inputs():
    rng = np.random.default_rng(11)
    V = (300.0 * rng.standard_normal(SHAPE)).astype(np.float32)
    psi0 = (rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)).astype(np.complex64)
    return V, psi0


@pytest.fixture(scope="module")
def oracle_results(inputs):
    from oracle import fdtd3d_c
    V, psi0 = inputs
    out = {}
    for eq in ("klein_gordon", "schrodinger"):
        p = _params(eq, "f64")
        d = p.oracle_dict()
        st = psi0.astype(np.complex128)
        if eq == "klein_gordon":
            st = (st, (-1j * p.e_sim * psi0).astype(np.complex64).astype(np.complex128))
        st = fdtd3d_c.run(st, V.astype(np.float64), NSTEPS, d, threads=16)
        out[eq] = np.concatenate([np.ravel(x) for x in (st if isinstance(st, tuple) else (st,))])
    return out


@pytest.mark.parametrize("eq,precision,tol", [
    ("klein_gordon", "f32", 3e-6), ("klein_gordon", "f32_kahan", 3e-6),
    ("klein_gordon", "f64", 1e-12), ("schrodinger", "f32", 3e-6), ("schrodinger", "f64", 1e-12)])
def test_full_grid_against_oracle(eq, precision, tol, inputs, oracle_results):
    from wideanglepropagation_amd.engine import FDTD3D
    p = _params(eq, precision)
    print(f"case full grid {eq} {precision} ({p.resolved_schedule()}) {SHAPE}", flush=True)
    V, psi0 = inputs
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64) if eq == "klein_gordon" else None
    sim = FDTD3D(p, V=V, psi0=psi0, pi0=pi0)
    try:
        sim.step(NSTEPS)
        st = sim.get_state(np.complex128)
    finally:
        sim.close()
    got = np.concatenate([np.ravel(x) for x in (st if isinstance(st, tuple) else (st,))])
    want = oracle_results[eq]
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"rel L2 vs oracle: {rel:.3e}", flush=True)
    assert rel < tol, rel


def test_512x512x256_grid_against_oracle():
    # 512 x 512 x 256 (67 M voxels): 64-plane z chunks, ten full 32-column groups per XCD
    # order (the grid above has a narrower last group)
    from oracle import fdtd3d_c
    from wideanglepropagation_amd.engine import FDTD3D
    from wideanglepropagation_amd.physics import klein_gordon_params
    from wideanglepropagation_amd.workloads import centred_extent
    shape = (256, 512, 512)
    p = klein_gordon_params(0.05, centred_extent(shape, 0.05, 0.005), shape, pml_thick=0.08,
                            k_source=18, precision="f32")
    print(f"case 512x512x256 grid klein_gordon f32 ({p.resolved_schedule()}) {shape}", flush=True)
    rng = np.random.default_rng(12)
    V = (300.0 * rng.standard_normal(shape)).astype(np.float32)
    psi0 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64)
    sim = FDTD3D(p, V=V, psi0=psi0, pi0=pi0)
    try:
        sim.step(2)
        got = np.concatenate([np.ravel(x) for x in sim.get_state(np.complex128)])
    finally:
        sim.close()
    st = fdtd3d_c.run((psi0.astype(np.complex128), pi0.astype(np.complex128)),
                      V.astype(np.float64), 2, p.with_(precision="f64").oracle_dict(), threads=16)
    want = np.concatenate([np.ravel(x) for x in st])
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"rel L2 vs oracle: {rel:.3e}", flush=True)
    assert rel < 3e-6, rel
