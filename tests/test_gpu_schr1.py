"""The one-pass Schrodinger schedule (schr1_pass: the four Horner levels of an RK4 step in
one walk over the planes) against the FP64 CPU oracle and against the two-pass schedule.

Tolerances: relative L2 2e-6 against the oracle after up to 20 steps (float32 arithmetic,
as test_gpu_parity), 1e-12 in float64 (rounding only; its workgroup has 11 row waves, 5
interior rows); against the two-pass float32 schedule, which evaluates the same levels
(outside the PML with the z term as a second difference), 1e-6.
"""
import numpy as np
import pytest

from oracle import fdtd3d

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.engine import FDTD3D
    return FDTD3D


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _params(shape, **kw):
    from wideanglepropagation_amd.physics import schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent
    ext = centred_extent(shape, kw.pop("spacing", 0.05))
    return schrodinger_params(0.05, ext, shape, pml_thick=kw.pop("pml_thick", 0.15),
                              k_source=kw.pop("k_source", 4),
                              precision=kw.pop("precision", "f32"), **kw)


def _case(p, rng, vscale=2e3):
    V = (vscale * rng.standard_normal(p.shape)).astype(np.float32)
    psi = (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)) * 0.5
    return V, psi


def _gpu_run(p, V, psi, nsteps, profile=None):
    FDTD3D = _gpu()
    sim = FDTD3D(p, V=V.astype(np.float64) if p.precision == "f64" else V, source_profile=profile)
    sim.set_state(psi)
    sim.step(nsteps)
    out = sim.get_state(np.complex128)
    sim.close()
    return out


def _oracle(p, V, psi, nsteps, profile=None):
    return fdtd3d.run(psi, V.astype(np.float64), nsteps, p.oracle_dict(src_profile=profile),
                      f32_time=(p.time_mode == "f32"))


# (nz, ny, nx) and z chunk: partial x tiles (70, 125, 130 columns), grids narrower than the
# tile's 8 interior rows (3, 5), one-plane and short chunks (chunk edges inside the source
# band and next to the frozen planes), a single chunk (0: the cost model's choice)
@pytest.mark.parametrize("shape,tz", [((9, 5, 70), 1), ((9, 5, 70), 3), ((11, 13, 130), 100),
                                      ((7, 3, 2), 0), ((12, 40, 125), 5), ((10, 3, 61), 2),
                                      ((40, 26, 64), 0), ((33, 17, 64), 7)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_schr1_matches_oracle_ragged(shape, tz, prec, rng):
    print(f"case schr_1pass {prec} {shape} tz={tz}", flush=True)
    p = _params(shape, tile_z=tz, pml_thick=0.1, k_source=2, schedule="schr_1pass",
                precision=prec)
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 8)
    want = _oracle(p, V, psi, 8)
    assert rel_l2(got, want) < (1e-12 if prec == "f64" else 2e-6)


# the steady-state step (all levels outside the PML, no source plane, no clamped plane)
# against the general one: a source band in the middle of the steady-state range (it splits
# the range), a thin and a thick PML, a PML on the end planes only, chunk edges inside the
# range; whole-grid chunks (the cost model picks 2-plane chunks on these small grids, too
# short for a steady-state step)
@pytest.mark.parametrize("shape,tz,k_source,pml", [((60, 20, 64), 60, 30, 0.15),
                                                   ((60, 20, 64), 16, 29, 0.15),
                                                   ((70, 12, 64), 70, 8, 0.6),
                                                   ((50, 9, 40), 13, 3, 0.025)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_schr1_steady_state_range(shape, tz, k_source, pml, prec, rng):
    print(f"case schr_1pass {prec} steady state {shape} tz={tz} k_source={k_source} pml={pml}",
          flush=True)
    p = _params(shape, tile_z=tz, pml_thick=pml, k_source=k_source, schedule="schr_1pass",
                precision=prec)
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 10)
    assert rel_l2(got, _oracle(p, V, psi, 10)) < (1e-12 if prec == "f64" else 2e-6)


def test_schr1_matches_oracle_20_steps(rng):
    print("case schr_1pass 24x16x16 20 steps", flush=True)
    p = _params((24, 16, 16), schedule="schr_1pass")
    V, psi = _case(p, rng)
    assert rel_l2(_gpu_run(p, V, psi, 20), _oracle(p, V, psi, 20)) < 2e-6


def test_schr1_source_profile_and_reference_clock(rng):
    print("case schr_1pass source profile + f32 clock", flush=True)
    p = _params((20, 12, 16), reference_quantize=True, k_source=5, schedule="schr_1pass")
    assert p.time_mode == "f32"
    prof = np.exp(-np.linspace(-1, 1, 16)[None, :] ** 2 - np.linspace(-1, 1, 12)[:, None] ** 2)
    prof = (prof * np.exp(0.3j)).astype(np.complex128)
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 12, profile=prof)
    assert rel_l2(got, _oracle(p, V, psi, 12, profile=prof)) < 2e-6


@pytest.mark.parametrize("shape,tz", [((40, 48, 96), 0), ((64, 40, 128), 9)])
def test_schr1_equals_two_pass(shape, tz, rng):
    # the same Horner levels with the same arithmetic as schr2_pass
    print(f"case schr_1pass vs schr_fused {shape} tz={tz}", flush=True)
    p1 = _params(shape, tile_z=tz, schedule="schr_1pass")
    p2 = p1.with_(schedule="schr_fused", tile_z=0)
    V, psi = _case(p1, rng)
    a = _gpu_run(p1, V, psi, 10)
    b = _gpu_run(p2, V, psi, 10)
    err = rel_l2(a, b)
    print(f"rel L2 one pass vs two passes: {err:.3e}", flush=True)
    assert err < 1e-6


def test_schr1_gaussian_64_200_steps():
    # the 64^3 wavepacket over 200 steps: as close to the FP64 run as the two-pass schedule
    from wideanglepropagation_amd.workloads import schrodinger_gaussian_64
    print("case 64^3 wavepacket schr_1pass vs schr_fused vs f64", flush=True)
    FDTD3D = _gpu()
    p64, V, psi0, nsteps = schrodinger_gaussian_64("f64")
    out = {}
    for name, p in (("f64", p64), ("two", p64.with_(precision="f32", schedule="schr_fused")),
                    ("one", p64.with_(precision="f32", schedule="schr_1pass"))):
        sim = FDTD3D(p, V=V.astype(np.float64 if name == "f64" else np.float32))
        sim.set_state(psi0)
        sim.step(nsteps)
        out[name] = sim.get_state(np.complex128)
        sim.close()
    e1, e2 = rel_l2(out["one"], out["f64"]), rel_l2(out["two"], out["f64"])
    print(f"rel L2 vs f64: one pass {e1:.3e}, two passes {e2:.3e}", flush=True)
    assert e1 < 1e-5 and e1 < 1.5 * e2 + 1e-7


def test_schr1_bitwise_deterministic(rng):
    FDTD3D = _gpu()
    p = _params((40, 48, 96), schedule="schr_1pass")
    V, psi = _case(p, rng)
    outs = []
    for _ in range(2):
        sim = FDTD3D(p, V=V)
        sim.set_state(psi)
        sim.step(7)
        outs.append(sim.get_state(np.complex64))
        sim.close()
    assert np.array_equal(outs[0], outs[1])


def test_schr1_rejects_compensation_on_2d_grids():
    # f32_kahan runs one pass on 3D grids only
    FDTD3D = _gpu()
    from wideanglepropagation_amd._lib import WapError
    from wideanglepropagation_amd.workloads import centred_extent
    p = _params((12, 8, 8), schedule="schr_1pass")
    flat = p.with_(ny=1, extent=centred_extent((12, 1, 8), 0.05), precision="f32_kahan")
    with pytest.raises(WapError, match="one-pass Schrodinger schedule"):
        FDTD3D(flat)


def test_schr1_adjoint_equals_two_pass_adjoint(rng):
    # records of the one-pass schedule: the same float32 gradient as with the two-pass
    # records, to float32 rounding
    _gpu()
    from wideanglepropagation_amd.adjoint import exit_wave, far_field_intensity
    print("case schr_1pass adjoint", flush=True)
    shape = (22, 10, 12)
    p1 = _params(shape, schedule="schr_1pass")
    V0, psi0 = _case(p1, rng, vscale=300.0)
    grads, Es = [], []
    for p in (p1, p1.with_(schedule="schr_fused")):
        V = torch.tensor(V0, dtype=torch.float32, device="cuda", requires_grad=True)
        E = exit_wave(V, p, 25, p.nz - 5, samples={25: 1.0, 18: 0.5}, checkpoint_every=6,
                      psi0=psi0)
        torch.sum(far_field_intensity(E) ** 2).backward()
        grads.append(V.grad.detach().double().cpu().numpy())
        Es.append(E.detach().cpu().numpy())
    assert np.abs(grads[1]).max() > 0
    assert rel_l2(Es[0], Es[1]) < 1e-6
    rel = rel_l2(grads[0], grads[1])
    print(f"one-pass vs two-pass gradient rel L2 {rel:.3e}", flush=True)
    assert rel < 1e-4


def test_schr1_adjoint_workspace_has_no_pass_buffers():
    # the one-pass reverse step keeps no intermediate fields: the workspace is the two-pass
    # adjoint's less the three fields its first pass writes
    FDTD3D = _gpu()
    from wideanglepropagation_amd.adjoint import _Adjoint
    p1 = _params((22, 10, 12), schedule="schr_1pass")
    nb = {}
    for p in (p1, p1.with_(schedule="schr_fused")):
        sim = FDTD3D(p)
        nb[p.schedule] = (_Adjoint.workspace_bytes(sim, 1), _Adjoint.workspace_bytes(sim, 3))
        sim.close()
    field = -(-22 * 10 * 12 * 8 // 256) * 256
    assert nb["schr_fused"][0] - nb["schr_1pass"][0] == 3 * field
    assert nb["schr_fused"][1] - nb["schr_1pass"][1] == 3 * field


@pytest.mark.parametrize("shape,tz,k_source,pml", [((60, 20, 64), 60, 30, 0.15),
                                                   ((60, 20, 64), 16, 29, 0.15),
                                                   ((50, 9, 40), 13, 3, 0.025)])
def test_schr1_adjoint_steady_state_range(shape, tz, k_source, pml, rng):
    # the one-pass reverse step's steady-state variant (long chunks outside the PML, the
    # source band splitting the range, chunk edges inside it) against the two-pass adjoint
    _gpu()
    from wideanglepropagation_amd.adjoint import exit_wave, far_field_intensity
    print(f"case schr_1pass adjoint steady state {shape} tz={tz} k_source={k_source} "
          f"pml={pml}", flush=True)
    p1 = _params(shape, tile_z=tz, pml_thick=pml, k_source=k_source, schedule="schr_1pass")
    V0, psi0 = _case(p1, rng, vscale=300.0)
    grads = []
    for p in (p1, p1.with_(schedule="schr_fused", tile_z=0)):
        V = torch.tensor(V0, dtype=torch.float32, device="cuda", requires_grad=True)
        E = exit_wave(V, p, 12, p.nz - 4, samples={12: 1.0, 9: 0.5}, checkpoint_every=4,
                      psi0=psi0)
        torch.sum(far_field_intensity(E) ** 2).backward()
        grads.append(V.grad.detach().double().cpu().numpy())
    assert np.abs(grads[1]).max() > 0
    rel = rel_l2(grads[0], grads[1])
    print(f"one-pass vs two-pass gradient rel L2 {rel:.3e}", flush=True)
    assert rel < 1e-4


# the compensated one-pass step (f32_kahan): the oracle on ragged shapes and chunk edges,
# a steady-state range split by the source band (whose level-4 planes take the keep factor
# on the compensation), a source profile with the float32 clock
@pytest.mark.parametrize("shape,tz,k_source", [((9, 5, 70), 1, 2), ((11, 13, 130), 100, 2),
                                               ((12, 40, 125), 5, 2), ((60, 20, 64), 60, 30),
                                               ((60, 20, 64), 16, 29)])
def test_schr1_kahan_matches_oracle(shape, tz, k_source, rng):
    print(f"case schr_1pass f32_kahan {shape} tz={tz} k_source={k_source}", flush=True)
    p = _params(shape, tile_z=tz, pml_thick=0.15, k_source=k_source, schedule="schr_1pass")
    p = p.with_(precision="f32_kahan")
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 10)
    assert rel_l2(got, _oracle(p, V, psi, 10)) < 2e-6


def test_schr1_kahan_source_profile_and_reference_clock(rng):
    print("case schr_1pass f32_kahan source profile + f32 clock", flush=True)
    p = _params((20, 12, 16), reference_quantize=True, k_source=5, schedule="schr_1pass")
    p = p.with_(precision="f32_kahan")
    prof = np.exp(-np.linspace(-1, 1, 16)[None, :] ** 2 - np.linspace(-1, 1, 12)[:, None] ** 2)
    prof = (prof * np.exp(0.3j)).astype(np.complex128)
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 12, profile=prof)
    assert rel_l2(got, _oracle(p, V, psi, 12, profile=prof)) < 2e-6


def test_schr1_kahan_compensates(rng):
    # over many small steps the compensated one-pass step stays closer to float64 than the
    # plain float32 one, and as close as the compensated two-pass schedule (the same inputs;
    # errors against the float64 two-pass schedule)
    FDTD3D = _gpu()
    print("case schr_1pass kahan vs plain f32, 1000 small steps", flush=True)
    p = _params((48, 32, 64), pml_thick=0.1, k_source=4)
    p = p.with_(precision="f64", schedule="schr_fused", dt=p.dt / 50)
    V, psi = _case(p, rng)
    out = {}
    for name, q in (("f64", p), ("f32", p.with_(precision="f32", schedule="schr_1pass")),
                    ("kahan1", p.with_(precision="f32_kahan", schedule="schr_1pass")),
                    ("kahan2", p.with_(precision="f32_kahan", schedule="schr_fused"))):
        sim = FDTD3D(q, V=V.astype(np.float64) if name == "f64" else V)
        sim.set_state(psi)
        sim.step(1000)
        out[name] = sim.get_state(np.complex128)
        sim.close()
    err = {k: rel_l2(out[k], out["f64"]) for k in ("f32", "kahan1", "kahan2")}
    print(f"rel L2 vs f64 after 1000 steps: {err}", flush=True)
    assert err["kahan1"] < err["f32"] and err["kahan1"] < 2 * err["kahan2"] + 1e-9
