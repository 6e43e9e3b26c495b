"""HIP stepper (through the C ABI) against the FP64 CPU oracle on the same inputs.

Tolerances (relative L2 of each state component, psi and pi, against its own norm):
  f64        1e-12 after up to 200 steps (roundoff only)
  f32        2e-6 after 20 steps (float32 arithmetic vs float64)
  f32_kahan  2e-6 after 20 steps
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
    """The largest relative L2 error over the state's components: psi and, for
    Klein-Gordon, pi, each against its own norm (|pi| ~ E |psi| would hide an error in psi
    in a norm over both)."""
    a = a if isinstance(a, tuple) else (a,)
    b = b if isinstance(b, tuple) else (b,)
    return max(float(np.linalg.norm(np.ravel(x) - np.ravel(y)) / np.linalg.norm(np.ravel(y)))
               for x, y in zip(a, b))


def _params(shape, eq="schrodinger", precision="f64", **kw):
    from wideanglepropagation_amd.physics import schrodinger_params, klein_gordon_params
    from wideanglepropagation_amd.workloads import centred_extent
    lam = 0.05
    ext = centred_extent(shape, kw.pop("spacing", 0.05))
    pml = kw.pop("pml_thick", 3 * 0.05)
    ks = kw.pop("k_source", 4)
    if eq == "schrodinger":
        return schrodinger_params(lam, ext, shape, pml_thick=pml, k_source=ks,
                                  precision=precision, **kw)
    return klein_gordon_params(lam, ext, shape, pml_thick=pml, k_source=ks,
                               precision=precision, **kw)


def _random_case(p, rng, vscale=2e3):
    shape = p.shape
    V = vscale * rng.standard_normal(shape)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.5
    if p.equation == "klein_gordon":
        pi = -1j * p.e_sim * psi * (1 + 0.01 * rng.standard_normal(shape))
        return V, (psi, pi)
    return V, psi


def _run_both(p, V, st0, nsteps, profile=None):
    FDTD3D = _gpu()
    if p.precision != "f64":
        V = V.astype(np.float32)
    sim = FDTD3D(p, V=V, source_profile=profile)
    if isinstance(st0, tuple):
        sim.set_state(st0[0], st0[1])
    else:
        sim.set_state(st0)
    sim.step(nsteps)
    got = sim.get_state(np.complex128)
    sim.close()
    want = fdtd3d.run(st0, V.astype(np.float64), nsteps, p.oracle_dict(src_profile=profile),
                      f32_time=(p.time_mode == "f32"))
    return got, want


@pytest.mark.parametrize("eq,sched", [("schrodinger", "auto"), ("schrodinger", "stages"),
                                       ("klein_gordon", "kg_2pass"),
                                       ("klein_gordon", "stages"), ("klein_gordon", "kg_fused")])
def test_f64_matches_oracle(eq, sched, rng):
    print(f"case f64 {eq} {sched} 16x16x24", flush=True)
    p = _params((24, 16, 16), eq, schedule=sched)
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 20)
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("eq,sched", [("schrodinger", "auto"), ("schrodinger", "stages"),
                                       ("klein_gordon", "kg_2pass"),
                                       ("klein_gordon", "kg_fused")])
@pytest.mark.parametrize("shape,ty,tz", [((9, 5, 70), 1, 1), ((9, 5, 70), 2, 3),
                                          ((11, 13, 130), 8, 100), ((7, 3, 2), 4, 0),
                                          ((12, 40, 125), 16, 5), ((10, 3, 61), 4, 2)])
def test_f64_ragged_shapes_and_tiles(shape, ty, tz, eq, sched, rng):
    print(f"case ragged {eq} {sched} {shape} ty={ty} tz={tz}", flush=True)
    if ty == 16 and sched != "kg_fused":
        ty = 8
    p = _params(shape, eq, tile_y=ty, tile_z=tz, pml_thick=0.1, k_source=2, schedule=sched)
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 6)
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("eq,sched", [("schrodinger", "stages"), ("schrodinger", "schr_fused"),
                                       ("klein_gordon", "kg_2pass"),
                                       ("klein_gordon", "stages"), ("klein_gordon", "kg_fused")])
@pytest.mark.parametrize("prec", ["f32", "f32_kahan"])
def test_f32_matches_oracle(eq, sched, prec, rng):
    print(f"case {prec} {eq} {sched}", flush=True)
    p = _params((24, 16, 16), eq, precision=prec, schedule=sched)
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 20)
    assert rel_l2(got, want) < 2e-6


@pytest.mark.parametrize("shape,ty,tz", [((9, 5, 70), 4, 1), ((9, 5, 70), 8, 3),
                                          ((11, 13, 130), 16, 100), ((7, 3, 2), 4, 0),
                                          ((12, 40, 125), 16, 5), ((10, 3, 61), 8, 2),
                                          ((40, 26, 64), 16, 0)])
@pytest.mark.parametrize("prec", ["f32", "f64", "f32_kahan"])
def test_schr_fused_ragged_shapes_and_tiles(shape, ty, tz, prec, rng):
    # the 2-pass Schrodinger schedule against the FP64 oracle on partial tiles, grids
    # narrower than a tile, short z chunks, and both Horner passes' frozen planes (the
    # source band sits on planes 2..4, next to the frozen plane 0), in every precision
    print(f"case schr_fused {prec} {shape} ty={ty} tz={tz}", flush=True)
    p = _params(shape, "schrodinger", precision=prec, tile_y=ty, tile_z=tz, pml_thick=0.1,
                k_source=2, schedule="schr_fused")
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 8)
    assert rel_l2(got, want) < (1e-12 if prec == "f64" else 2e-6)


@pytest.mark.parametrize("shape,ty,tz", [((9, 5, 70), 4, 1), ((11, 13, 130), 16, 100),
                                          ((12, 40, 125), 16, 5), ((10, 3, 61), 8, 2),
                                          ((40, 26, 64), 16, 0)])
def test_kg_fused_kahan_ragged_shapes_and_tiles(shape, ty, tz, rng):
    # the compensated fused pass (compensation loaded one plane ahead, updated in place)
    # on partial tiles, short z chunks (chunk edges) and the source band
    print(f"case kg_fused f32_kahan {shape} ty={ty} tz={tz}", flush=True)
    p = _params(shape, "klein_gordon", precision="f32_kahan", tile_y=ty, tile_z=tz,
                pml_thick=0.1, k_source=2, schedule="kg_fused")
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 8)
    assert rel_l2(got, want) < 2e-6


def test_kg_fused_kahan_compensates(rng):
    # over many steps the compensated fused pass stays closer to float64 than the plain
    # float32 one (the same inputs; errors against the float64 fused pass)
    from wideanglepropagation_amd.engine import FDTD3D
    print("case kg_fused kahan vs plain f32, 1000 small steps", flush=True)
    p = _params((48, 32, 64), "klein_gordon", precision="f64", pml_thick=0.1, k_source=4)
    p = p.with_(dt=p.dt / 50)   # small increments: the state accumulation's rounding dominates
    V, st0 = _random_case(p, rng)
    V = V.astype(np.float32).astype(np.float64)   # the same potential in every precision
    out = {}
    for prec in ("f64", "f32", "f32_kahan"):
        q = p.with_(precision=prec)
        assert q.resolved_schedule() == "kg_fused"
        sim = FDTD3D(q, V=V)
        sim.set_state(st0[0], st0[1])
        sim.step(1000)
        st = sim.get_state(np.complex128)
        sim.close()
        out[prec] = np.concatenate([np.ravel(x) for x in st])
    err = {k: rel_l2(out[k], out["f64"]) for k in ("f32", "f32_kahan")}
    print(f"rel L2 vs f64 after 1000 steps: {err}", flush=True)
    assert err["f32_kahan"] < err["f32"]


def test_kg_fused_is_the_default_in_every_precision():
    p = _params((24, 16, 16), "klein_gordon", precision="f32")
    assert p.resolved_schedule() == "kg_fused"
    assert p.with_(precision="f64").resolved_schedule() == "kg_fused"
    assert p.with_(precision="f32_kahan").resolved_schedule() == "kg_fused"


def test_schr_fused_is_the_default_in_every_precision():
    # every precision: the four Horner levels in one pass (two passes on request)
    p = _params((24, 16, 16), "schrodinger", precision="f32")
    assert p.resolved_schedule() == "schr_1pass"
    assert p.with_(precision="f64").resolved_schedule() == "schr_1pass"
    assert p.with_(precision="f32_kahan").resolved_schedule() == "schr_1pass"
    # the 2D rule (zero stage inputs on the end rows): one pass too (its end-plane factors)
    assert p.with_(end_planes="zero").resolved_schedule() == "schr_1pass"
    assert p.with_(end_planes="zero", precision="f64").resolved_schedule() == "schr_1pass"


def test_schr_fused_kahan_compensates(rng):
    # the compensated Schrodinger fused pass stays closer to float64 than the plain float32
    # one over many small steps (the same inputs; errors against the float64 fused pass)
    from wideanglepropagation_amd.engine import FDTD3D
    print("case schr_fused kahan vs plain f32, 1000 small steps", flush=True)
    p = _params((48, 32, 64), "schrodinger", precision="f64", pml_thick=0.1, k_source=4)
    p = p.with_(dt=p.dt / 50)
    V, st0 = _random_case(p, rng)
    V = V.astype(np.float32).astype(np.float64)
    out = {}
    for prec in ("f64", "f32", "f32_kahan"):
        q = p.with_(precision=prec, schedule="schr_fused")
        sim = FDTD3D(q, V=V)
        sim.set_state(st0)
        sim.step(1000)
        out[prec] = sim.get_state(np.complex128)
        sim.close()
    err = {k: rel_l2(out[k], out["f64"]) for k in ("f32", "f32_kahan")}
    print(f"rel L2 vs f64 after 1000 steps: {err}", flush=True)
    assert err["f32_kahan"] < err["f32"]


def test_source_profile_and_reference_clock(rng):
    print("case source profile + f32 clock", flush=True)
    p = _params((20, 12, 16), reference_quantize=True, k_source=5)
    assert p.time_mode == "f32"
    prof = np.exp(-np.linspace(-1, 1, 16)[None, :] ** 2 - np.linspace(-1, 1, 12)[:, None] ** 2)
    prof = prof * np.exp(0.3j)
    V, st0 = _random_case(p, rng)
    got, want = _run_both(p, V, st0, 12, profile=prof.astype(np.complex128))
    assert rel_l2(got, want) < 1e-12


def test_schrodinger_gaussian_64_f64_200_steps():
    from wideanglepropagation_amd.workloads import schrodinger_gaussian_64
    print("case 64^3 wavepacket f64 200 steps", flush=True)
    p, V, psi0, nsteps = schrodinger_gaussian_64("f64")
    got, want = _run_both(p, V, psi0, nsteps)
    assert rel_l2(got, want) < 1e-11


@pytest.mark.parametrize("eq", ["schrodinger", "klein_gordon"])
def test_f32_is_bitwise_deterministic(eq, rng):
    FDTD3D = _gpu()
    p = _params((40, 48, 96), eq, precision="f32")
    V, st0 = _random_case(p, rng)
    outs = []
    for _ in range(2):
        sim = FDTD3D(p, V=V.astype(np.float32))
        sim.set_state(*(st0 if isinstance(st0, tuple) else (st0,)))
        sim.step(10)
        outs.append(sim.get_state(np.complex64))
        sim.close()
    a, b = outs
    for x, y in zip(a if isinstance(a, tuple) else (a,), b if isinstance(b, tuple) else (b,)):
        assert np.array_equal(x, y)


def test_norm_plane_and_accumulator(rng):
    FDTD3D = _gpu()
    p = _params((12, 10, 20), precision="f32_kahan")
    V, st0 = _random_case(p, rng)
    sim = FDTD3D(p, V=V.astype(np.float32), psi0=st0)
    sim.step(3)
    st = sim.get_state(np.complex128)
    assert sim.norm2() == pytest.approx(float(np.sum(np.abs(st) ** 2)), rel=1e-12)
    assert np.array_equal(sim.get_plane(5), st[5])
    acc = torch.zeros((10, 20), dtype=torch.complex128, device="cuda")
    sim.accumulate_plane(5, 0.7, acc)
    sim.accumulate_plane(5, 0.7, acc)
    assert np.allclose(acc.cpu().numpy(), 2 * st[5] * np.exp(0.7j), rtol=1e-14, atol=1e-14)
    sim.close()


def test_errors_are_reported_not_fatal():
    FDTD3D = _gpu()
    from wideanglepropagation_amd._lib import WapError
    p = _params((12, 8, 8))
    sim = FDTD3D(p)
    with pytest.raises(WapError, match="potential not set"):
        sim.step(1)
    sim.close()
    with pytest.raises(WapError, match="grid must be"):
        FDTD3D(p.with_(nz=2))
    with pytest.raises(WapError, match="source plane"):
        FDTD3D(p.with_(src_k=99))
    with pytest.raises(WapError, match="for the Klein-Gordon equation"):
        FDTD3D(p.with_(schedule="kg_2pass"))
    with pytest.raises(WapError, match="for the Klein-Gordon equation"):
        FDTD3D(p.with_(schedule="kg_fused"))
    with pytest.raises(WapError, match="Schrodinger fused schedule"):
        FDTD3D(p.with_(schedule="schr_fused", end_planes="zero"))
    with pytest.raises(WapError, match="Schrodinger fused schedule"):
        FDTD3D(p.with_(schedule="schr_fused", precision="f32", equation="klein_gordon"))
    with pytest.raises(ValueError):
        FDTD3D(p, V=np.zeros(7))


def test_schr_fused_precision_64cubed_200_steps():
    # the 64^3 wavepacket over 200 steps: the Horner-form 2-pass schedule in float32 is
    # as close to the FP64 run as one pass per stage in float32
    from wideanglepropagation_amd.workloads import schrodinger_gaussian_64
    FDTD3D = _gpu()
    print("case 64^3 wavepacket f32 schr_fused vs stages vs f64", flush=True)
    p64, V, psi0, nsteps = schrodinger_gaussian_64("f64")
    out = {}
    for name, p in (("f64", p64), ("stages", p64.with_(precision="f32", schedule="stages")),
                    ("fused", p64.with_(precision="f32", schedule="schr_fused"))):
        sim = FDTD3D(p, V=V.astype(np.float64 if name == "f64" else np.float32))
        sim.set_state(psi0)
        sim.step(nsteps)
        out[name] = sim.get_state(np.complex128)
        sim.close()
    e_st = rel_l2(out["stages"], out["f64"])
    e_fu = rel_l2(out["fused"], out["f64"])
    print(f"rel L2 vs f64: stages {e_st:.3e}, fused {e_fu:.3e}", flush=True)
    assert e_fu < 1e-5 and e_fu < 1.5 * e_st + 1e-7
