"""The fused passes on a grid whose step fields outgrow the Infinity Cache.

Above 256 MB of step fields (4 x the state's size) the fused passes store their results
with non-temporal stores: a second copy of the Klein-Gordon pass body, a per-store
policy in the Schrodinger passes and in both adjoints.  The other parity tests use
small grids, which take the cached copy.  Here a 128 x 128 x 544 grid (8.9 M voxels,
4 x 71 MB of step fields) runs the float32 fused schedules against the float64 GPU
schedules (the oracle pins those at small sizes), and the float32 fused adjoints
against the float64 adjoints.  Tolerances as for the small grids: 3e-6 relative L2 for
a few steps of the state, 2e-3 for the float32 gradient.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (544, 128, 128)


def _params(eq, precision):
    from wideanglepropagation_amd.physics import klein_gordon_params, schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent
    mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
    return mk(0.05, centred_extent(SHAPE, 0.05, 0.005), SHAPE, pml_thick=0.08, k_source=18,
              precision=precision)


def _inputs(rng, p):
    V = (300.0 * rng.standard_normal(SHAPE)).astype(np.float32)
    psi0 = (rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)).astype(np.complex64)
    pi0 = (-1j * p.e_sim * psi0).astype(np.complex64) if p.equation == "klein_gordon" else None
    return V, psi0, pi0


def test_grid_is_above_the_cache_rule():
    nz, ny, nx = SHAPE
    assert nx * ny * nz * 8 * 4 > 256 << 20


@pytest.mark.parametrize("eq", ["klein_gordon", "schrodinger"])
def test_fused_forward_with_streaming_stores_matches_f64(eq, rng):
    from wideanglepropagation_amd.engine import FDTD3D
    print(f"case stream-store forward {eq} {SHAPE}", flush=True)
    p32, p64 = _params(eq, "f32"), _params(eq, "f64")
    assert p32.resolved_schedule() in ("kg_fused", "schr_1pass")
    V, psi0, pi0 = _inputs(rng, p32)
    out = []
    for p in (p32, p64):
        sim = FDTD3D(p, V=V, psi0=psi0, pi0=pi0)
        sim.step(8)
        st = sim.get_state(np.complex128)
        sim.close()
        out.append(np.concatenate([np.ravel(x) for x in (st if isinstance(st, tuple) else (st,))]))
    rel = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert rel < 3e-6, rel


@pytest.mark.parametrize("eq", ["klein_gordon", "schrodinger"])
def test_fused_adjoint_with_streaming_stores_matches_f64(eq, rng):
    from wideanglepropagation_amd.adjoint import exit_wave, far_field_intensity
    print(f"case stream-store adjoint {eq} {SHAPE}", flush=True)
    p32, p64 = _params(eq, "f32"), _params(eq, "f64")
    V, psi0, pi0 = _inputs(rng, p32)
    grads = []
    for p, dt in ((p64, torch.float64), (p32, torch.float32)):
        Vt = torch.tensor(V, dtype=dt, device="cuda", requires_grad=True)
        E = exit_wave(Vt, p, 12, p.nz - 20, psi0=psi0, pi0=pi0)
        torch.sum(far_field_intensity(E) ** 2).backward()
        grads.append(Vt.grad.detach().double().cpu().numpy())
    rel = np.linalg.norm(grads[1] - grads[0]) / np.linalg.norm(grads[0])
    assert np.isfinite(rel) and rel < 2e-3, rel
