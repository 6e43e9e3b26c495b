"""The one-pass Schrodinger schedule on 2D grids (ny = 1: schr2d_pass, one wave per
56-column x tile and z chunk, x neighbours exchanged between lanes) and with the end-plane
rule of the reference's 2D program (stage inputs zero on the end rows), against the FP64 CPU
oracle, which evaluates that rule stage by stage (``oracle/fdtd3d.py``, pinned by the
sparse-matrix brute force in ``tests/test_oracle.py``).

Tolerances: relative L2 1e-12 in float64 (rounding only); 2e-6 in float32 after up to 20
steps (as test_gpu_parity); against the stage schedule in float32, 2e-6.
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
    ext = centred_extent(shape, 0.05)
    return schrodinger_params(0.05, ext, shape, pml_thick=kw.pop("pml_thick", 0.15),
                              k_source=kw.pop("k_source", 3),
                              precision=kw.pop("precision", "f32"),
                              end_planes=kw.pop("end_planes", "zero"), **kw)


def _case(p, rng, vscale=2e3):
    V = (vscale * rng.standard_normal(p.shape)).astype(np.float32)
    psi = (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)) * 0.5
    return V, psi


def _gpu_run(p, V, psi, nsteps, profile=None):
    FDTD3D = _gpu()
    sim = FDTD3D(p, V=V if p.precision != "f64" else V.astype(np.float64), source_profile=profile)
    sim.set_state(psi)
    sim.step(nsteps)
    out = sim.get_state(np.complex128)
    sim.close()
    return out


def _oracle(p, V, psi, nsteps, profile=None):
    return fdtd3d.run(psi, V.astype(np.float64), nsteps, p.oracle_dict(src_profile=profile),
                      f32_time=(p.time_mode == "f32"))


def test_2d_grids_default_to_one_pass():
    p = _params((30, 1, 64))
    assert p.resolved_schedule() == "schr_1pass"
    assert p.with_(precision="f64").resolved_schedule() == "schr_1pass"
    assert p.with_(end_planes="frozen").resolved_schedule() == "schr_1pass"
    FDTD3D = _gpu()
    sim = FDTD3D(p)
    ty, tz, groups = sim.launch_shape()
    sim.close()
    assert ty == 1 and groups == 2 * ((30 + tz - 1) // tz)   # 64 columns: two 56-column tiles


# (nz, 1, nx) and z chunk: one tile narrower than the grid's 56 columns (32), exactly one
# tile (56), a one-column partial tile (57), several tiles with a partial last one (130,
# 200), one-plane and short chunks, chunks cut inside the source band, a whole-grid chunk
@pytest.mark.parametrize("shape,tz", [((20, 1, 32), 1), ((20, 1, 56), 2), ((21, 1, 57), 5),
                                      ((33, 1, 130), 0), ((40, 1, 200), 7), ((24, 1, 64), 24)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_2d_one_pass_matches_oracle(shape, tz, prec, rng):
    print(f"case schr2d {prec} {shape} tz={tz}", flush=True)
    p = _params(shape, tile_z=tz, precision=prec, schedule="schr_1pass")
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 10)
    assert rel_l2(got, _oracle(p, V, psi, 10)) < (1e-12 if prec == "f64" else 2e-6)


# long grids: chunks whose first steps run the steady-state ramp (outside the PML and the
# source band), blocks mixing steady-state and general steps, a chunk edge inside the band
@pytest.mark.parametrize("tz,k_source", [(8, 60), (13, 57), (0, 33)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_2d_one_pass_long_grid(tz, k_source, prec, rng):
    print(f"case schr2d long {prec} tz={tz} k_source={k_source}", flush=True)
    p = _params((120, 1, 100), tile_z=tz, precision=prec, pml_thick=0.3, k_source=k_source,
                schedule="schr_1pass")
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 10)
    assert rel_l2(got, _oracle(p, V, psi, 10)) < (1e-12 if prec == "f64" else 2e-6)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_2d_one_pass_frozen_ends(prec, rng):
    # a 2D grid with the 3D rule (every stage input keeps b on the end planes)
    print(f"case schr2d frozen {prec}", flush=True)
    p = _params((26, 1, 90), precision=prec, end_planes="frozen", schedule="schr_1pass")
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 12)
    assert rel_l2(got, _oracle(p, V, psi, 12)) < (1e-12 if prec == "f64" else 2e-6)


def test_2d_one_pass_source_profile_and_reference_clock(rng):
    print("case schr2d source profile + f32 clock", flush=True)
    p = _params((30, 1, 100), reference_quantize=True, k_source=6, schedule="schr_1pass")
    assert p.time_mode == "f32"
    prof = np.exp(-np.linspace(-1, 1, 100) ** 2)[None, :] * np.exp(0.4j)
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 20, profile=prof)
    assert rel_l2(got, _oracle(p, V, psi, 20, profile=prof)) < 2e-6


def test_2d_one_pass_equals_stages(rng):
    print("case schr2d vs stages", flush=True)
    p = _params((64, 1, 300), pml_thick=0.4, k_source=12, schedule="schr_1pass")
    V, psi = _case(p, rng)
    a = _gpu_run(p, V, psi, 15)
    b = _gpu_run(p.with_(schedule="stages"), V, psi, 15)
    err = rel_l2(a, b)
    print(f"rel L2 one pass vs stages: {err:.3e}", flush=True)
    assert err < 2e-6


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_3d_zero_end_planes_one_pass(prec, rng):
    # the end-plane factors of the 3D one-pass kernel (zero end planes on a 3D grid)
    print(f"case schr_1pass 3D zero end planes {prec}", flush=True)
    p = _params((18, 12, 70), end_planes="zero", schedule="schr_1pass", tile_z=5, precision=prec)
    assert p.with_(schedule="auto").resolved_schedule() == "schr_1pass"
    V, psi = _case(p, rng)
    got = _gpu_run(p, V, psi, 10)
    assert rel_l2(got, _oracle(p, V, psi, 10)) < (1e-12 if prec == "f64" else 2e-6)


def test_zero_end_planes_reject_compensation():
    # the 2D rule is the reference's 2D program's, which has no compensated precision
    FDTD3D = _gpu()
    from wideanglepropagation_amd._lib import WapError
    p = _params((18, 12, 70), end_planes="zero", precision="f32_kahan")
    with pytest.raises(WapError, match="zero end planes are for the f32/f64 Schrodinger"):
        FDTD3D(p)


def test_2d_batch_matches_oracle(rng):
    # batched 2D grids run the 3D one-pass kernel with the 2D end-plane factors
    _gpu()
    from wideanglepropagation_amd.batched import BatchedFDTD
    from wideanglepropagation_amd.engine import FDTD3D
    print("case batched 2D one pass", flush=True)
    ps = [_params((22, 1, 40), k_source=3), _params((30, 1, 90), k_source=5)]
    cases = [_case(p, rng) for p in ps]
    sims = []
    for p, (V, psi) in zip(ps, cases):
        s = FDTD3D(p, V=V)
        s.set_state(psi)
        sims.append(s)
    b = BatchedFDTD(sims)
    b.step(9)
    for p, s, (V, psi) in zip(ps, sims, cases):
        assert rel_l2(s.get_state(np.complex128), _oracle(p, V, psi, 9)) < 2e-6
    b.close()
    for s in sims:
        s.close()


def test_2d_float64_batch_rejected(rng):
    FDTD3D = _gpu()
    from wideanglepropagation_amd._lib import WapError
    from wideanglepropagation_amd.batched import BatchedFDTD
    p = _params((22, 1, 40), precision="f64")
    s = FDTD3D(p, V=np.zeros(p.shape))
    with pytest.raises(WapError, match="batched float64 one-pass Schrodinger grids are 3D"):
        BatchedFDTD([s])
    s.close()


def test_2d_adjoint_records_from_the_2d_step(rng):
    # a 2D grid with frozen end planes (the adjoint's rule): records made by schr2d_pass,
    # reverse steps by schr_adj1_pass on the one-row grid, against the two-pass pair
    _gpu()
    from wideanglepropagation_amd.adjoint import exit_wave, far_field_intensity
    print("case schr2d adjoint", flush=True)
    p1 = _params((40, 1, 64), end_planes="frozen", k_source=4, schedule="schr_1pass")
    V0, psi0 = _case(p1, rng, vscale=300.0)
    grads = []
    for p in (p1, p1.with_(schedule="schr_fused")):
        V = torch.tensor(V0, dtype=torch.float32, device="cuda", requires_grad=True)
        E = exit_wave(V, p, 20, p.nz - 5, samples={20: 1.0, 14: 0.5}, checkpoint_every=5,
                      psi0=psi0)
        torch.sum(far_field_intensity(E) ** 2).backward()
        grads.append(V.grad.detach().double().cpu().numpy())
    assert np.abs(grads[1]).max() > 0
    rel = rel_l2(grads[0], grads[1])
    print(f"2D one-pass vs two-pass gradient rel L2 {rel:.3e}", flush=True)
    assert rel < 1e-4


@pytest.mark.parametrize("prec,tol", [("f32", 3e-6), ("f64", 1e-12)])
def test_2d_reference_grid_against_c_oracle(prec, tol, rng):
    # the reference's fdtd.cu grid, NX 1024 x NY 1398 (here nx x nz), through the drop-in
    # 2D parameters (atomic units, float32 clock and constants), 40 steps of the default
    # schedule (the one-pass 2D step, 4-plane chunks) against the FP64 C + OpenMP oracle
    from oracle import fdtd3d_c
    from wideanglepropagation_amd.fdtd_solver import RelativisticSolver
    FDTD3D = _gpu()
    s = RelativisticSolver()
    p = s._params2d((1398, 1024), (-40.0, 40.0, -60.0, 60.0), 0.05, 4.0, 80.0, 120,
                    {"precision": prec})
    assert p.resolved_schedule() == "schr_1pass" and (p.nz, p.ny, p.nx) == (1398, 1, 1024)
    print(f"case schr2d reference grid {prec}", flush=True)
    V = (0.05 * rng.standard_normal(p.shape)).astype(np.float32)
    psi = ((rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)) * 0.5)
    sim = FDTD3D(p, V=V if prec == "f32" else V.astype(np.float64))
    sim.set_state(psi)
    sim.step(40)
    got = sim.get_state(np.complex128)
    sim.close()
    want = fdtd3d_c.run(psi.astype(np.complex128), V.astype(np.float64), 40, p.oracle_dict(),
                        f32_time=(p.time_mode == "f32"), threads=16)
    err = rel_l2(got, want)
    # the state moves ~1 % in 40 steps: the increments are compared too (float32 rounds the
    # state to complex64 every step, ~4e-7 of it over 40 steps, ~3e-5 of the increment)
    inc = rel_l2(got - psi, want - psi)
    print(f"rel L2 vs C oracle: state {err:.3e}, increment {inc:.3e}", flush=True)
    assert err < tol and inc < (1e-4 if prec == "f32" else 1e-10)
