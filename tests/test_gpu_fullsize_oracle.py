"""The HIP path against the oracle at the benchmark's full grid.

256 x 256 x 512 (nx, ny, nz; 33.5 M voxels): the grid of bench.py's default run, whose
fields outgrow the Infinity Cache, so the fused passes take their non-temporal-store
copies and run 1280 workgroups over 16 z chunks.  A random state goes through 3 steps of
each default schedule (Klein-Gordon fused in float32, float32 + Kahan and float64;
Schrodinger fused in float32, stage passes in float64).  The result is compared with
the FP64 C + OpenMP restatement of the oracle's step (oracle/fdtd3d_omp.c, itself pinned
to the NumPy oracle in tests/test_oracle_c.py) on the same inputs.  Tolerances:
float64 1e-12, float32 3e-6 relative L2 of each state component against its own norm (as
for the small grids).  The Klein-Gordon fused pass is also checked on 512 x 512 x 256
(2 steps) in float32, float32 + Kahan and float64.
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


def _rel(got, want):
    """Largest relative L2 over the state's components (psi, and pi for Klein-Gordon),
    each against its own norm."""
    return max(float(np.linalg.norm(g - w) / np.linalg.norm(w)) for g, w in zip(got, want))


def _components(st):
    return [np.ravel(x) for x in (st if isinstance(st, tuple) else (st,))]


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
        out[eq] = _components(st)
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
    rel = _rel(_components(st), oracle_results[eq])
    print(f"rel L2 vs oracle (worst component): {rel:.3e}", flush=True)
    assert rel < tol, rel


SHAPE_512 = (256, 512, 512)


@pytest.fixture(scope="module")
def inputs_512():
    rng = np.random.default_rng(12)
    V = (300.0 * rng.standard_normal(SHAPE_512)).astype(np.float32)
    psi0 = (rng.standard_normal(SHAPE_512) + 1j * rng.standard_normal(SHAPE_512)).astype(np.complex64)
    return V, psi0


def _params_512(precision):
    from wideanglepropagation_amd.physics import klein_gordon_params
    from wideanglepropagation_amd.workloads import centred_extent
    return klein_gordon_params(0.05, centred_extent(SHAPE_512, 0.05, 0.005), SHAPE_512,
                               pml_thick=0.08, k_source=18, precision=precision)


@pytest.fixture(scope="module")
def oracle_512(inputs_512):
    from oracle import fdtd3d_c
    V, psi0 = inputs_512
    p = _params_512("f64")
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64)
    st = fdtd3d_c.run((psi0.astype(np.complex128), pi0.astype(np.complex128)),
                      V.astype(np.float64), 2, p.oracle_dict(), threads=16)
    return _components(st)


@pytest.mark.parametrize("precision,tol", [("f32", 3e-6), ("f32_kahan", 3e-6), ("f64", 1e-12)])
def test_512x512x256_grid_against_oracle(precision, tol, inputs_512, oracle_512):
    # the precision study's grid, 512 x 512 x 256 (67 M voxels): 64-plane z chunks, ten
    # full 32-column groups per XCD order (the grid above has a narrower last group);
    # the three precisions the study compares
    from wideanglepropagation_amd.engine import FDTD3D
    p = _params_512(precision)
    print(f"case 512x512x256 grid klein_gordon {precision} ({p.resolved_schedule()})",
          flush=True)
    V, psi0 = inputs_512
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64)
    sim = FDTD3D(p, V=V if precision != "f64" else V.astype(np.float64), psi0=psi0, pi0=pi0)
    try:
        sim.step(2)
        got = _components(sim.get_state(np.complex128))
    finally:
        sim.close()
    rel = _rel(got, oracle_512)
    print(f"rel L2 vs oracle (worst component): {rel:.3e}", flush=True)
    assert rel < tol, rel
