"""The C + OpenMP restatement of the oracle's step (oracle/fdtd3d_omp.c, bench.py's CPU
baseline) against the NumPy oracle: Schrodinger and Klein-Gordon, PML / sponge, the soft
source with and without a lateral profile, the 2D zero-end rule, several thread counts.
Both are FP64 and sum the RK4 terms in different orders: agreement to 1e-12 relative L2.
"""
import numpy as np
import pytest

from oracle import fdtd3d, fdtd3d_c


def _p(eq, shape, rng, **kw):
    nz, ny, nx = shape
    p = {"nx": nx, "ny": ny, "nz": nz, "xmin": -0.4, "xmax": 0.4, "ymin": -0.3, "ymax": 0.3,
         "zmin": -0.6, "zmax": 0.7, "dt": 2e-6, "A": 2.738256, "e_sim": 171870.6,
         "k_beam": 250.53233, "equation": eq, "hbar_c": 1973.2698, "rest_energy": 510998.95,
         "pml_thick": 0.2, "pml_sigma0": 80.0, "pml_theta": 0.785, "pml_power": 3.0,
         "kg_gamma_max": 4e4, "src_enabled": True, "src_k": 3, "src_band": 3,
         "src_alpha": 0.05, "src_profile": None}
    if eq == "klein_gordon":
        p["dt"] = 2e-7
    if ny == 1:
        p["ymin"] = p["ymax"] = 0.0
    p.update(kw)
    return p


def _state(p, rng):
    sh = (p["nz"], p["ny"], p["nx"])
    psi = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    if p["equation"] == "schrodinger":
        return psi
    return (psi, -1j * p["e_sim"] * psi + rng.standard_normal(sh))


def _rel(a, b):
    a = np.concatenate([np.ravel(x) for x in (a if isinstance(a, tuple) else (a,))])
    b = np.concatenate([np.ravel(x) for x in (b if isinstance(b, tuple) else (b,))])
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("eq", ["schrodinger", "klein_gordon"])
@pytest.mark.parametrize("case", ["plain", "profile", "no_source", "2d_zero_ends"])
def test_c_oracle_matches_numpy(eq, case, rng):
    shape = (20, 1, 24) if case == "2d_zero_ends" else (18, 10, 12)
    kw = {}
    if case == "profile":
        kw["src_profile"] = (rng.standard_normal(shape[1:]) +
                             1j * rng.standard_normal(shape[1:]))
    if case == "no_source":
        kw["src_enabled"] = False
    if case == "2d_zero_ends":
        kw["end_planes"] = "zero"
    p = _p(eq, shape, rng, **kw)
    V = 300.0 * rng.standard_normal(shape)
    st = _state(p, rng)
    want = fdtd3d.run(st, V, 6, p, t0=1e-5, step0=3)
    for threads in (1, 3):
        got = fdtd3d_c.run(st, V, 6, p, t0=1e-5, step0=3, threads=threads)
        assert _rel(got, want) < 1e-12, (threads, _rel(got, want))


def test_c_oracle_reference_clock_and_input_untouched(rng):
    p = _p("schrodinger", (14, 8, 8), rng)
    V = 200.0 * rng.standard_normal((14, 8, 8))
    st = _state(p, rng)
    keep = st.copy()
    want = fdtd3d.run(st, V, 4, p, step0=7, f32_time=True)
    got = fdtd3d_c.run(st, V, 4, p, step0=7, f32_time=True)
    assert _rel(got, want) < 1e-12
    assert np.array_equal(st, keep)


def test_c_oracle_catches_a_dropped_term(rng):
    # a planted mistake in the NumPy oracle's sponge (gamma -> 0) must fail the check above
    p = _p("klein_gordon", (18, 10, 12), rng)
    V = 300.0 * rng.standard_normal((18, 10, 12))
    st = _state(p, rng)
    got = fdtd3d_c.run(st, V, 4, p)
    bad = dict(p, kg_gamma_max=0.0)
    assert _rel(got, fdtd3d.run(st, V, 4, bad)) > 1e-9
