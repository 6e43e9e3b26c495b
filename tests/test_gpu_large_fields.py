"""Fused passes on a grid whose fields exceed 2 GiB.

The fused kernels address each field through buffer descriptors based at the first plane
of their workgroup's chunk window, so the 32-bit offsets cover a window, not the field.
272 x 1024 x 1024 (nz, ny, nx; 285 M voxels): a complex64 field is 2.28 GB, beyond the
2 GiB that whole-field offsets could reach, and the chunks are capped so that a window
(the chunk and its z halo) stays under 2 GiB.  A random state goes through 2 steps of the
default float32 schedules (Klein-Gordon fused, Schrodinger one pass) and is compared with
the FP64 C + OpenMP restatement of the oracle's step (oracle/fdtd3d_omp.c, pinned to the
NumPy oracle in tests/test_oracle_c.py).  Tolerance: float32 3e-6 relative L2 of each
state component against its own norm, as for the smaller grids.  The test needs ~60 GB of
host memory for the oracle's complex128 arrays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (272, 1024, 1024)
NSTEPS = 2


def _params(eq, precision):
    from wideanglepropagation_amd.physics import klein_gordon_params, schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent
    mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
    return mk(0.05, centred_extent(SHAPE, 0.05, 0.005), SHAPE, pml_thick=0.08, k_source=18,
              precision=precision)


@pytest.fixture(scope="module")
def inputs():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rng = np.random.default_rng(13)
    V = (300.0 * rng.standard_normal(SHAPE, dtype=np.float32)).astype(np.float32)
    psi0 = np.empty(SHAPE, np.complex64)
    psi0.real = rng.standard_normal(SHAPE, dtype=np.float32)
    psi0.imag = rng.standard_normal(SHAPE, dtype=np.float32)
    return V, psi0


def _rel(got, want):
    return max(float(np.linalg.norm(g - w) / np.linalg.norm(w)) for g, w in zip(got, want))


@pytest.mark.parametrize("eq", ["klein_gordon", "schrodinger"])
def test_fields_over_2gib_against_oracle(eq, inputs):
    from oracle import fdtd3d_c
    from wideanglepropagation_amd.engine import FDTD3D
    V, psi0 = inputs
    p = _params(eq, "f32")
    assert p.voxels * 8 > 2 ** 31 and p.resolved_schedule() in ("kg_fused", "schr_1pass")
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64) if eq == "klein_gordon" else None
    print(f"case {eq} f32 {SHAPE} ({p.resolved_schedule()}), {p.voxels * 8 / 2**30:.2f} GiB "
          "per field", flush=True)
    sim = FDTD3D(p, V=V, psi0=psi0, pi0=pi0)
    try:
        print(f"launch shape {sim.launch_shape()}", flush=True)
        sim.step(NSTEPS)
        got = sim.get_state(np.complex64)
    finally:
        sim.close()
    got = [np.ravel(x) for x in (got if isinstance(got, tuple) else (got,))]
    d = _params(eq, "f64").oracle_dict()
    st = psi0.astype(np.complex128)
    if eq == "klein_gordon":
        st = (st, pi0.astype(np.complex128))
    want = fdtd3d_c.run(st, V.astype(np.float64), NSTEPS, d, threads=16)
    want = [np.ravel(x) for x in (want if isinstance(want, tuple) else (want,))]
    rel = _rel([g.astype(np.complex128) for g in got], want)
    print(f"rel L2 vs oracle (worst component): {rel:.3e}", flush=True)
    assert rel < 3e-6, rel
