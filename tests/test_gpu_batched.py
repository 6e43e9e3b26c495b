"""Batched launches (one launch per pass over many uneven grids) and their hipGraph.

A batched step runs the same per-column device code as a single-grid step;
the two kernels are compiled separately, so the compiler may contract a few
multiply-adds differently: every instance must match its own single-grid run
to rounding (relative L2 below 1e-6 for float32, 1e-13 for float64).  Graph
replays run the batched kernels themselves and must match the eager batched
steps bit for bit on every replay.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _cases(rng, eq, prec, n=5):
    from wideanglepropagation_amd.physics import (schrodinger_params, klein_gordon_params,
                                                  wavelength_from_kinetic_energy)
    from wideanglepropagation_amd.workloads import centred_extent
    shapes = [(20, 16, 16), (24, 12, 70), (18, 33, 20), (30, 8, 64), (21, 17, 129)][:n]
    out = []
    for i, shape in enumerate(shapes):
        lam = wavelength_from_kinetic_energy(60e3 + 50e3 * i)
        ext = centred_extent(shape, 0.05)
        mk = schrodinger_params if eq == "schrodinger" else klein_gordon_params
        p = mk(lam, ext, shape, pml_thick=0.15, k_source=4, precision=prec, tile_z=3 + i)
        V = (500 * rng.standard_normal(shape)).astype(np.float64 if prec == "f64" else np.float32)
        yy, xx = np.meshgrid(np.linspace(-1, 1, shape[1]), np.linspace(-1, 1, shape[2]),
                             indexing="ij")
        prof = np.exp(-((xx - 0.2 * i) ** 2 + yy ** 2) / 0.3) * np.exp(-0.5j * (xx ** 2 + yy ** 2))
        out.append((p, V, prof))
    return out


def _make(cases):
    from wideanglepropagation_amd.engine import FDTD3D
    return [FDTD3D(p, V=V, source_profile=prof) for p, V, prof in cases]


@pytest.mark.parametrize("eq,prec", [("schrodinger", "f32"), ("klein_gordon", "f32"),
                                     ("klein_gordon", "f32_kahan"), ("schrodinger", "f64"),
                                     ("schrodinger", "f32_kahan")])
def test_batched_step_equals_single_grids(eq, prec, rng):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.batched import BatchedFDTD
    print(f"case batched {eq} {prec}", flush=True)
    cases = _cases(rng, eq, prec)
    single = _make(cases)
    for s in single:
        s.step(7)
    want = [s.get_state(np.complex128) for s in single]
    batch_sims = _make(cases)
    b = BatchedFDTD(batch_sims, tile_y=4)
    b.step(3)
    b.step(4)
    b.synchronize()
    tol = 1e-13 if prec == "f64" else 1e-6
    for s, w in zip(batch_sims, want):
        got = s.get_state(np.complex128)
        for g, x in zip(got if isinstance(got, tuple) else (got,),
                        w if isinstance(w, tuple) else (w,)):
            assert np.linalg.norm(g - x) <= tol * np.linalg.norm(x)
    b.close()
    for s in single + batch_sims:
        s.close()


def test_graph_replay_matches_eager_steps(rng):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.batched import BatchedFDTD
    print("case graph replay", flush=True)
    cases = _cases(rng, "schrodinger", "f32") + _cases(rng, "klein_gordon", "f32", n=3)
    eager = _make(cases)
    eb = BatchedFDTD(eager)
    graph = _make(cases)
    gb = BatchedFDTD(graph)
    gb.capture(4)
    for rep in range(3):
        eb.step(4)
        ms = gb.replay(timed=True)
        assert ms > 0
        # eager work between replays
        x = torch.ones(1000, device="cuda").sum().item()
        assert x == 1000
        gb.synchronize()
        for e, g in zip(eager, graph):
            a, c = e.get_state(np.complex64), g.get_state(np.complex64)
            if isinstance(a, tuple):
                assert all(np.array_equal(u, v) for u, v in zip(a, c)), rep
            else:
                assert np.array_equal(a, c), rep
    eb.close()
    gb.close()
    for s in eager + graph:
        s.close()


def test_replay_after_longer_eager_steps_matches_eager(rng):
    # eager steps longer than the captured graph regrow the eager phase table; the graph
    # keeps its own table, so a replay afterwards still runs the right source phases
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.batched import BatchedFDTD
    print("case capture(4) step(8) replay()", flush=True)
    cases = _cases(rng, "schrodinger", "f32", n=3) + _cases(rng, "klein_gordon", "f32", n=2)
    eager = _make(cases)
    eb = BatchedFDTD(eager)
    graph = _make(cases)
    gb = BatchedFDTD(graph)
    gb.capture(4)
    gb.step(8)
    eb.step(8)
    gb.replay()
    eb.step(4)
    gb.replay()
    eb.step(4)
    gb.synchronize()
    eb.synchronize()
    for e, g in zip(eager, graph):
        a, c = e.get_state(np.complex64), g.get_state(np.complex64)
        for u, v in zip(a if isinstance(a, tuple) else (a,), c if isinstance(c, tuple) else (c,)):
            assert np.array_equal(u, v)
    eb.close()
    gb.close()
    for s in eager + graph:
        s.close()


def test_batch_rejects_mixed_precision(rng):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd._lib import WapError, lib
    import ctypes
    cases = _cases(rng, "schrodinger", "f32", n=1) + _cases(rng, "schrodinger", "f64", n=1)
    sims = _make(cases)
    arr = (ctypes.c_void_p * 2)(*[s.handle.value for s in sims])
    b = ctypes.c_void_p()
    rc = lib().wap_batch_create(arr, 2, 0, ctypes.byref(b))
    assert rc != 0 and b"share equation" in lib().wap_last_error()
    for s in sims:
        s.close()


@pytest.mark.parametrize("eq", ["schrodinger", "klein_gordon"])
def test_fused_batch_odd_steps_and_odd_graph(eq, rng):
    # fused batches ping-pong the states: odd eager steps and replays of an odd-length
    # graph leave the state in the other buffer, which the handles must then point to
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wideanglepropagation_amd.batched import BatchedFDTD
    print(f"case fused batch odd steps {eq}", flush=True)
    cases = _cases(rng, eq, "f32", n=4)
    assert all(p.resolved_schedule() in ("kg_fused", "schr_1pass") for p, _, _ in cases)
    single = _make(cases)
    for s in single:
        s.step(9)
    want = [s.get_state(np.complex64) for s in single]
    sims = _make(cases)
    b = BatchedFDTD(sims)
    b.step(3)
    b.capture(3)
    b.replay()
    b.replay()
    b.synchronize()
    for s, w in zip(sims, want):
        got = s.get_state(np.complex64)
        for g, x in zip(got if isinstance(got, tuple) else (got,),
                        w if isinstance(w, tuple) else (w,)):
            assert np.linalg.norm(g - x) <= 1e-6 * np.linalg.norm(x)
    # the handles keep stepping on their own after the batch
    for s in single + sims:
        s.step(2)
    for s, t in zip(single, sims):
        a, c = s.get_state(np.complex64), t.get_state(np.complex64)
        for g, x in zip(c if isinstance(c, tuple) else (c,), a if isinstance(a, tuple) else (a,)):
            assert np.linalg.norm(g - x) <= 1e-6 * np.linalg.norm(x)
    b.close()
    for s in single + sims:
        s.close()


def test_autotune_single_grid_returns_a_measured_candidate(tmp_path, monkeypatch):
    from wideanglepropagation_amd import autotune
    from wideanglepropagation_amd.physics import klein_gordon_params
    from wideanglepropagation_amd.workloads import centred_extent
    monkeypatch.setenv("WAP_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    shape = (40, 26, 64)
    p = klein_gordon_params(0.025, centred_extent(shape, 0.05), shape, pml_thick=0.15,
                            k_source=4, precision="f32")
    ty, tz, ms, table = autotune.tune(p, tile_ys=(8, 16), tile_zs=(8, 16), steps=2, warmup=1)
    assert len(table) == 4 and all(m > 0 for _, _, m in table)
    assert (ty, tz, ms) == min(table, key=lambda r: r[2])
    # the cached answer comes back without timing again
    assert autotune.tune(p, tile_ys=(8, 16), tile_zs=(8, 16))[:2] == (ty, tz)


def test_tune_batch_picks_a_chunk_per_class():
    from wideanglepropagation_amd import autotune
    from wideanglepropagation_amd.physics import klein_gordon_params, schrodinger_params
    from wideanglepropagation_amd.workloads import centred_extent
    ps = []
    for i, shape in enumerate([(40, 26, 64), (48, 20, 40), (40, 26, 64), (48, 20, 40)]):
        mk = schrodinger_params if i % 2 else klein_gordon_params
        ps.append(mk(0.025, centred_extent(shape, 0.05), shape, pml_thick=0.15, k_source=4,
                     precision="f32"))
    choice, ms, table = autotune.tune_batch(ps, tile_zs=(8, 16), tile_ys=(8, 16), steps=2,
                                            warmup=1)
    assert set(choice) == {autotune._class_of(p) for p in ps}
    assert {tz for _, tz in choice.values()} <= {8, 16}
    assert {ty for ty, _ in choice.values()} <= {8, 16}
    # the grids of one launch group share its waves
    for c in choice:
        assert all(choice[d][0] == choice[c][0] for d in choice if d[:3] == c[:3])
    assert ms == min(m for _, m in table) and ms > 0
    # 2 uniform chunks, one alternative waves per group whose shape is tunable (the
    # one-pass Schrodinger group's is fixed), one alternative chunk per class
    tunable = {c[:3] for c in choice if c[2] != "schr_1pass"}
    assert len(table) == 2 + len(tunable) + len(choice)


def test_batched_study_quick_runs_finite():
    from wideanglepropagation_amd.studies import batched_study
    r = batched_study(quick=True)
    assert r["all_finite"] and r["gvoxel_steps_per_s"] > 0
    # the float32 Schrodinger classes run the one-pass step, whose workgroup is 15 waves
    assert r["autotuned"] and all(
        v["tile_z"] in (16, 32, 64)
        and (v["waves"] == 15 if k.startswith("schrodinger/") else v["waves"] in (8, 16))
        for k, v in r["autotuned"].items())
