"""Pins the CPU oracle (oracle/) against golden values and mathematics.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign, index or axis, an in-place stage) fails it:
  * reference-held golden scalars (notebook outputs, checked-in macros);
  * brute force: one oracle step == P4(dt M) applied to the injected state,
    with M assembled voxel by voxel as a scipy.sparse matrix;
  * closed-form discrete dispersion in a laterally periodic box
    (Schrodinger and Klein-Gordon);
  * PML profile closed forms and absorption, norm behaviour.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fdtd3d, extract, physics


# --------------------------------------------------------------- constants
def test_au_constants_match_notebook_outputs():
    # FDTDmethod/fdtd_3D_plane_wave.ipynb:50 prints "66.49836952842307 2211.016574969353"
    _, p_inf, e_sim = physics.calculate_physics_params(0.05)
    assert p_inf == pytest.approx(66.49836952842307, rel=1e-15)
    assert e_sim == pytest.approx(2211.016574969353, rel=1e-15)
    # FDTDmethod/fdtd_1D_baseline_comparison.ipynb:79-80: k0 132.9967, E_sim 8844.0663 (a.u.)
    _, p_inf, e_sim = physics.calculate_physics_params(0.025)
    assert round(p_inf, 4) == 132.9967
    assert round(e_sim, 4) == 8844.0663
    assert physics.calculate_physics_params(0.0) == (0, 0, 0)


def test_angstrom_constants_match_checked_in_macros():
    # fdtd_3d.cu:44,48 carry the lambda = 0.05 A values written by _update_macros
    q = physics.quantized_constants(0.05, (-1, 1, -1, 1, -1.6, 1.5811))
    c = physics.angstrom_constants(0.05)
    assert "%.6f" % c["A"] == "3.427719"          # the text _update_macros writes
    assert "%.6f" % c["P_INF"] == "125.663706"
    assert q["A"] == float(np.float32(3.427719))   # the float literal nvcc parses
    assert q["P_INF"] == float(np.float32(125.663706))
    # fdtd_3D_plane_wave.ipynb:219 prints E_sim=54128.367188 (float32 product, fdtd_3d.cu:316)
    assert "%.6f" % q["E_sim"] == "54128.367188"
    # ... and phase per frame 108.2567 rad at dt = 2e-6, 1000 steps/frame (:301)
    assert round(q["E_sim"] * 1000 * float(np.float32(2e-6)), 4) == 108.2567
    # k_beam = sqrtf(E_sim / A) recovers P_INF to float precision (fdtd_3d.cu:139)
    assert q["k_beam"] == pytest.approx(125.663706, rel=2e-7)
    assert q["dt"] == float(np.float32(5e-7))


def test_angstrom_closed_form_200kev():
    lam = 0.0250793
    c = physics.angstrom_constants(lam)
    # E_total = mc^2 + 200 keV to the wavelength's 6 digits
    assert c["E_total"] - physics.REST_ENERGY_EV == pytest.approx(200e3, rel=1e-4)
    assert c["A"] == pytest.approx(physics.HBAR_C_EV_A ** 2 / (2 * c["E_total"]), rel=1e-15)
    assert c["A"] == pytest.approx(2.738256, rel=1e-5)


# ----------------------------------------------------------------- helpers
def _params(nx, ny, nz, eq="schrodinger", pml_cells=2, src=True, k_src=3, **kw):
    dx = 0.05
    ext = (-(nx - 1) * dx / 2, (nx - 1) * dx / 2, -(ny - 1) * dx / 2, (ny - 1) * dx / 2,
           -(nz - 1) * dx / 2, (nz - 1) * dx / 2)
    p = {
        "nx": nx, "ny": ny, "nz": nz,
        "xmin": ext[0], "xmax": ext[1], "ymin": ext[2], "ymax": ext[3],
        "zmin": ext[4], "zmax": ext[5],
        "dt": 5e-7, "A": 2.738256, "e_sim": 171870.6, "k_beam": 250.53,
        "equation": eq, "hbar_c": 1973.2698, "rest_energy": 510998.95,
        "pml_thick": pml_cells * dx, "pml_sigma0": 80.0, "pml_theta": 0.785, "pml_power": 3.0,
        "kg_gamma_max": 2.0e5, "src_enabled": src, "src_k": k_src, "src_band": 3,
        "src_alpha": 0.05, "src_profile": None,
    }
    if eq == "klein_gordon":
        p["e_sim"] = 710998.95
    p.update(kw)
    return p


def _a_profile_ref(p):
    """a(z) from the reference formula, written out per plane (fdtd_3d.cu:96-123)."""
    out = []
    for k in range(p["nz"]):
        z = p["zmin"] + k * (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0)
        rb = max((p["zmin"] + p["pml_thick"]) - z, 0.0) / p["pml_thick"]
        rt = max(z - (p["zmax"] - p["pml_thick"]), 0.0) / p["pml_thick"]
        r = max(rb, rt)
        sig = 0.0 if r <= 0.0 else p["pml_sigma0"] * r ** p["pml_power"]
        if sig <= 0.0 or p["equation"] != "schrodinger":
            out.append(1.0 + 0j)
        else:
            out.append(1.0 / complex(1.0 + sig * math.cos(p["pml_theta"]),
                                     sig * math.sin(p["pml_theta"])))
    return out


def _lap_matrix(p):
    """Assemble the 3D operator voxel by voxel (rows of frozen planes are empty)."""
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    dx, dy, dz = fdtd3d.spacing(p)
    a = _a_profile_ref(p)
    rows, cols, vals = [], [], []

    def ix(i, j, k):
        return k * nx * ny + j * nx + i

    for k in range(1, nz - 1):
        ap = 0.5 * (a[k] + a[k + 1])
        am = 0.5 * (a[k - 1] + a[k])
        for j in range(ny):
            for i in range(nx):
                r = ix(i, j, k)
                ent = [
                    (ix((i + 1) % nx, j, k), 1 / dx ** 2), (ix((i - 1) % nx, j, k), 1 / dx ** 2),
                    (ix(i, (j + 1) % ny, k), 1 / dy ** 2), (ix(i, (j - 1) % ny, k), 1 / dy ** 2),
                    (r, -2 / dx ** 2 - 2 / dy ** 2),
                    (ix(i, j, k + 1), a[k] * ap / dz ** 2),
                    (ix(i, j, k - 1), a[k] * am / dz ** 2),
                    (r, -a[k] * (ap + am) / dz ** 2),
                ]
                for c, v in ent:
                    rows.append(r)
                    cols.append(c)
                    vals.append(v)
    n = nx * ny * nz
    return sp.csr_matrix((np.array(vals, dtype=np.complex128), (rows, cols)), shape=(n, n))


def _interior_mask(p):
    m = np.zeros((p["nz"], p["ny"], p["nx"]), dtype=bool)
    m[1:-1] = True
    return m.ravel()


def _p4(M, b, dt):
    out = b.copy()
    term = b.copy()
    for j in range(1, 5):
        term = dt * (M @ term) / j
        out = out + term
    return out


# ------------------------------------------------------------- brute force
@pytest.mark.parametrize("src", [False, True])
def test_schrodinger_step_equals_p4_of_sparse_H(src, rng):
    p = _params(5, 4, 9, src=src, pml_cells=3, k_src=2)
    shape = (p["nz"], p["ny"], p["nx"])
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    V = 3e4 * rng.standard_normal(shape)
    L = _lap_matrix(p)
    H = -p["A"] * L + sp.diags(V.ravel() * _interior_mask(p))
    t = 7 * p["dt"]
    b = fdtd3d.inject(psi, t, p)
    want = _p4(-1j * H, b.ravel(), p["dt"]).reshape(shape)
    got = fdtd3d.rk4_step(psi, V, t, p)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13
    # frozen planes keep the injected base value
    assert np.array_equal(got[0], b[0]) and np.array_equal(got[-1], b[-1])


def test_klein_gordon_step_equals_p4_of_sparse_system(rng):
    p = _params(4, 5, 8, eq="klein_gordon", pml_cells=3, k_src=2)
    shape = (p["nz"], p["ny"], p["nx"])
    n = np.prod(shape)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    pi = 7e5 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    V = 1e3 * rng.standard_normal(shape)
    mask = _interior_mask(p)
    L = _lap_matrix(p)
    g = fdtd3d.kg_gamma(p)[:, None, None] * np.ones(shape)
    c2, m2 = p["hbar_c"] ** 2, p["rest_energy"] ** 2
    D = sp.diags(mask.astype(float))
    top = sp.hstack([sp.csr_matrix((n, n)), D])
    bot = sp.hstack([c2 * L - sp.diags((m2 - V.ravel() ** 2) * mask),
                     sp.diags((-2j * V.ravel() - 2 * g.ravel()) * mask)])
    M = sp.vstack([top, bot]).tocsr()
    t = 3 * p["dt"]
    b = fdtd3d.inject((psi, pi), t, p)
    want = _p4(M, np.concatenate([b[0].ravel(), b[1].ravel()]), p["dt"])
    got = fdtd3d.rk4_step((psi, pi), V, t, p)
    w_psi, w_pi = want[:n].reshape(shape), want[n:].reshape(shape)
    assert np.max(np.abs(got[0] - w_psi)) / np.max(np.abs(w_psi)) < 1e-12
    assert np.max(np.abs(got[1] - w_pi)) / np.max(np.abs(w_pi)) < 1e-12


def test_2d_zero_end_rows_step_equals_projected_rk4_of_sparse_H(rng):
    # fdtd.cu: stage outputs are 0 on the end rows (189-198), the final update
    # skips them (272); a 2D grid is ny = 1 with A = 0.5
    p = _params(6, 1, 9, src=True, pml_cells=3, k_src=2, A=0.5, end_planes="zero")
    shape = (p["nz"], 1, p["nx"])
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    V = 3e4 * rng.standard_normal(shape)
    H = -0.5 * _lap_matrix(p) + sp.diags(V.ravel() * _interior_mask(p))
    M = -1j * H
    Z = sp.diags(_interior_mask(p).astype(float))
    t = 5 * p["dt"]
    dt = p["dt"]
    b = fdtd3d.inject(psi, t, p).ravel()
    k1 = M @ b
    k2 = M @ (Z @ (b + 0.5 * dt * k1))
    k3 = M @ (Z @ (b + 0.5 * dt * k2))
    k4 = M @ (Z @ (b + dt * k3))
    want = (b + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)).reshape(shape)
    got = fdtd3d.rk4_step(psi, V, t, p)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13
    assert np.array_equal(got[0], b.reshape(shape)[0])
    # the end rows matter: the frozen-plane rule differs next to them by far
    # more than the tolerance above (it enters at second order in dt)
    frozen = fdtd3d.rk4_step(psi, V, t, dict(p, end_planes="frozen"))
    assert np.max(np.abs(frozen[1] - got[1])) > 1e-9 * np.max(np.abs(got[1]))


def test_2d_grid_is_the_y_uniform_3d_grid(rng):
    # ny = 1 has no y term: a y-uniform state on ny = 2 steps identically
    p1 = _params(6, 1, 10, k_src=3, A=0.5, end_planes="zero")
    p2 = dict(p1, ny=2, ymin=-0.025, ymax=0.025)
    psi = rng.standard_normal((10, 1, 6)) + 1j * rng.standard_normal((10, 1, 6))
    V = 1e4 * rng.standard_normal((10, 1, 6))
    a = fdtd3d.run(psi, V, 3, p1)
    b = fdtd3d.run(np.repeat(psi, 2, axis=1), np.repeat(V, 2, axis=1), 3, p2)
    assert np.allclose(b[:, 0], a[:, 0], rtol=0, atol=1e-12 * np.max(np.abs(a)))
    assert np.allclose(b[:, 1], a[:, 0], rtol=0, atol=1e-12 * np.max(np.abs(a)))


def test_zero_and_frozen_end_rules_agree_on_zero_end_planes(rng):
    # with zero end planes and no source there, k and b are 0 on the ends, so
    # both rules write the same stage outputs
    p = _params(5, 1, 9, k_src=3, A=0.5)
    psi = rng.standard_normal((9, 1, 5)) + 1j * rng.standard_normal((9, 1, 5))
    psi[[0, -1]] = 0
    V = 1e4 * rng.standard_normal((9, 1, 5))
    a = fdtd3d.run(psi, V, 4, dict(p, end_planes="zero"))
    b = fdtd3d.run(psi, V, 4, dict(p, end_planes="frozen"))
    assert np.array_equal(a, b)


def test_source_window_weights_and_phase():
    # fdtd_3d.cu:150-156: s = (k-K)/2 -> weights 0, 0.025, 0.05 on K, K+1, K+2
    p = _params(4, 4, 12, k_src=4)
    ks, ws = fdtd3d.source_weights(p)
    assert ks == [4, 5, 6]
    assert ws == pytest.approx([0.0, 0.025, 0.05], abs=1e-17)
    psi = np.zeros((12, 4, 4), complex)
    t = 11 * p["dt"]
    out = fdtd3d.inject(psi, t, p)
    z = fdtd3d.z_coords(p)
    for k, w in zip(ks, ws):
        ph = p["k_beam"] * (z[k] - z[4]) - p["e_sim"] * t
        assert np.allclose(out[k], w * np.exp(1j * ph), rtol=0, atol=1e-16)
    assert np.count_nonzero(out[:4]) == 0 and np.count_nonzero(out[7:]) == 0


# ---------------------------------------------------- closed-form physics
def _plane_wave(p, mx, my):
    nx, ny = p["nx"], p["ny"]
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    kx = 2 * np.pi * mx / nx
    ky = 2 * np.pi * my / ny
    lat = np.exp(1j * (kx * i + ky * j))
    dx, dy, _ = fdtd3d.spacing(p)
    K2 = 4 / dx ** 2 * np.sin(kx / 2) ** 2 + 4 / dy ** 2 * np.sin(ky / 2) ** 2
    return lat, K2


def test_schrodinger_discrete_dispersion():
    # z-uniform, laterally periodic mode of the discrete operator: one step multiplies the
    # planes more than 4 cells from the frozen ends by P4(-i dt (A K^2 + V0)) exactly.
    p = _params(8, 6, 16, src=False, pml_cells=0.5)
    lat, K2 = _plane_wave(p, 2, 1)
    V0 = -1234.5
    psi = np.broadcast_to(lat, (16, 6, 8)).astype(complex)
    V = np.full(psi.shape, V0)
    got = fdtd3d.rk4_step(psi, V, 0.0, p)
    lam = -1j * (p["A"] * K2 + V0) * p["dt"]
    amp = sum(lam ** j / math.factorial(j) for j in range(5))
    assert np.allclose(got[5:-5], amp * psi[5:-5], rtol=0, atol=1e-13)


@pytest.mark.parametrize("branch", [+1, -1])
def test_klein_gordon_discrete_dispersion(branch):
    p = _params(8, 6, 16, eq="klein_gordon", src=False, pml_cells=0.5)
    lat, K2 = _plane_wave(p, 1, 2)
    V0 = 350.0
    c2, m2 = p["hbar_c"] ** 2, p["rest_energy"] ** 2
    s = -1j * (V0 - branch * np.sqrt(c2 * K2 + m2))   # psi ~ e^{s t}
    psi = np.broadcast_to(lat, (16, 6, 8)).astype(complex)
    pi = s * psi
    V = np.full(psi.shape, V0)
    got = fdtd3d.rk4_step((psi, pi), V, 0.0, p)
    amp = sum((s * p["dt"]) ** j / math.factorial(j) for j in range(5))
    assert np.allclose(got[0][5:-5], amp * psi[5:-5], rtol=0, atol=1e-12)
    assert np.allclose(got[1][5:-5], amp * pi[5:-5], rtol=0, atol=1e-12 * abs(s))


def test_pml_profile_closed_form():
    p = _params(4, 4, 41, pml_cells=8)
    sig = fdtd3d.pml_sigma(p)
    assert sig[0] == pytest.approx(80.0) and sig[-1] == pytest.approx(80.0)
    assert sig[4] == pytest.approx(80.0 * 0.5 ** 3)       # half way into the layer
    assert np.all(sig[8:33] < 1e-30)                     # interior (edge planes round)
    a = fdtd3d.pml_a(p)
    assert a[0] == pytest.approx(1 / (1 + 80 * np.exp(0.785j)))
    assert np.allclose(a[8:33], 1.0, rtol=0, atol=1e-30)


def test_pml_absorbs_and_norm_never_grows():
    # a wavepacket moving towards +z into the layer loses its norm; with sigma = 0 the
    # RK4 amplification |P4(i theta)| <= 1 keeps it from growing
    p = _params(8, 8, 48, src=False, pml_cells=14, A=1.0, dt=1e-4)
    z = fdtd3d.z_coords(p)
    kz = 8.0
    env = np.exp(-((z - 0.6) / 0.15) ** 2) * np.exp(1j * kz * z)
    psi = np.broadcast_to(env[:, None, None], (48, 8, 8)).astype(complex)
    V = np.zeros(psi.shape)
    n0 = fdtd3d.norm2(psi)
    out = fdtd3d.run(psi, V, 3000, p)
    assert fdtd3d.norm2(out) < 0.2 * n0
    q = dict(p, pml_sigma0=0.0)
    out2 = fdtd3d.run(psi, V, 300, q)
    assert fdtd3d.norm2(out2) <= n0 * (1 + 1e-12)


def test_run_is_deterministic(rng):
    p = _params(6, 5, 10)
    V = 1e3 * rng.standard_normal((10, 5, 6))
    a = fdtd3d.run(fdtd3d.zero_state(p), V, 5, p)
    b = fdtd3d.run(fdtd3d.zero_state(p), V, 5, p)
    assert np.array_equal(a, b)


# -------------------------------------------------------------- extraction
def test_exit_wave_recipe_known_answer():
    e = 1234.5
    ny, nx, nz = 3, 4, 5
    amp = np.arange(ny * nx).reshape(ny, nx) + 1j
    snaps = []
    for m in range(10):
        t = np.float32(m * 1e-4)
        f = np.zeros((nz, ny, nx), complex)
        f[3] = amp * np.exp(-1j * e * float(t))
        f[1] = 2.0 * np.exp(-1j * e * float(t))
        snaps.append((t, f.real.astype(np.float32), f.imag.astype(np.float32)))
    ex, inc = extract.exit_wave(snaps, e, k_exit=3, k_monitor=1)
    assert inc == pytest.approx(4.0, rel=1e-6)
    assert np.allclose(ex, amp / 2.0, rtol=1e-5, atol=1e-5)
    ff = extract.far_field(np.ones((4, 4)))
    assert ff[2, 2] == pytest.approx(256.0) and ff.sum() == pytest.approx(256.0)
