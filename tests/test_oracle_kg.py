"""Pins the Klein-Gordon branches of the CPU oracle (oracle/fdtd3d.py) and its
float32 clock by mathematics that does not come from the oracle itself.

The reference has no Klein-Gordon stepper, so nothing it ships pins these
branches; each check below is a property of the KG equation

    d psi/dt = pi,   d pi/dt = c^2 Lap psi - (m^2 - V^2) psi - 2 i V pi - 2 gamma(z) pi

or of the reference's float32 time (``fdtd_3d.cu:365``), and each is shown to
fail on a mistake planted in the oracle (on the CPU):

  * the incident wave psi_inc = P e^{i(k(z - z_K) - w t)}, pi_inc = -i w psi_inc
    passes through ``inject`` unchanged (planted: +i w, a profile dropped
    from pi, a flipped k);
  * that incident wave, put on the discrete dispersion relation, steps by
    the closed-form RK4 amplification P4(-i w dt) (with the source on);
  * the sponge gamma(z) = gamma_max r^p: closed form, and one oracle step equals
    P4(dt M) with M assembled voxel by voxel from an independent gamma and an
    independent soft source (planted: r^(p+1));
  * the KG charge Q = sum Im(conj(psi) pi) + V |psi|^2 is conserved by the exact
    flow when gamma = 0, so RK4 keeps it to O(dt^5) over a fixed time (halving dt
    cuts the drift ~32x); planted: the 2 i V pi term dropped or sign-flipped;
  * with gamma > 0, dQ/dt = -2 sum gamma Im(conj(psi) pi): one small step moves Q
    by that amount (planted: gamma pi for 2 gamma pi, and the sign flipped);
  * the reference's forward-propagation invariant (``tests/test_fdtd_3d.py:312-323``)
    for the KG soft source: the demodulated phase rises along z with the
    discrete wavenumber (planted: +i w injection);
  * ``ref_time`` equals round_f32(round_f32(n) * round_f32(dt)) evaluated in exact
    rational arithmetic (the int * float product of ``fdtd_3d.cu:365``).
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fdtd3d

C2 = 1973.2698 ** 2          # (hbar c)^2 [eV^2 A^2]
M_REST = 510998.95            # m c^2 [eV]


def _kg(nx, ny, nz, dz, dx=0.05, **kw):
    ext = (-(nx - 1) * dx / 2, (nx - 1) * dx / 2, -(ny - 1) * dx / 2, (ny - 1) * dx / 2,
           -(nz - 1) * dz / 2, (nz - 1) * dz / 2)
    p = {
        "nx": nx, "ny": ny, "nz": nz,
        "xmin": ext[0], "xmax": ext[1], "ymin": ext[2], "ymax": ext[3],
        "zmin": ext[4], "zmax": ext[5],
        "dt": 5e-7, "A": 2.738256, "e_sim": 710998.95, "k_beam": 250.53,
        "equation": "klein_gordon", "hbar_c": 1973.2698, "rest_energy": M_REST,
        "pml_thick": 0.5 * dz, "pml_sigma0": 80.0, "pml_theta": 0.785, "pml_power": 3.0,
        "kg_gamma_max": 0.0, "src_enabled": True, "src_k": 4, "src_band": 3,
        "src_alpha": 0.05, "src_profile": None,
    }
    p.update(kw)
    return p


def _z(p):
    return [p["zmin"] + k * (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0) for k in range(p["nz"])]


def _window(p):
    """Soft-source weights a_k, written out per plane (fdtd_3d.cu:150-156)."""
    K, band = int(p["src_k"]), int(p["src_band"])
    return {K + b: p["src_alpha"] * 0.5 * (1.0 - math.cos(math.pi * b / (band - 1.0)))
            for b in range(band)}


def _incident(p, t, prof=None):
    """psi_inc and pi_inc = -i w psi_inc on every plane."""
    z = _z(p)
    zK = z[int(p["src_k"])]
    w, k = p["e_sim"], p["k_beam"]
    ph = np.array([complex(math.cos(k * (zz - zK) - w * t), math.sin(k * (zz - zK) - w * t))
                   for zz in z])
    lat = np.ones((p["ny"], p["nx"]), complex) if prof is None else prof
    psi = ph[:, None, None] * lat[None]
    return psi, -1j * w * psi


def _dev(a, b):
    return max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in zip(a, b))


# ------------------------------------------------------------ soft source
def _planted_plus_iw(state, t, p):
    """pi injected with +i w instead of -i w."""
    psi, pi = fdtd3d.inject(state, t, p)
    inc, _ = _incident(p, t, p.get("src_profile"))
    pi = pi.copy()
    for k, a in _window(p).items():
        pi[k] += a * 2j * p["e_sim"] * inc[k]
    return psi, pi


def _planted_profile_dropped_from_pi(state, t, p):
    """The lateral profile multiplies psi's injection but not pi's."""
    psi, pi = fdtd3d.inject(state, t, p)
    inc, _ = _incident(p, t, p.get("src_profile"))
    flat, _ = _incident(p, t, None)
    pi = pi.copy()
    for k, a in _window(p).items():
        pi[k] += a * (-1j * p["e_sim"]) * (flat[k] - inc[k])
    return psi, pi


def _planted_flipped_k(state, t, p):
    return fdtd3d.inject(state, t, dict(p, k_beam=-p["k_beam"]))


@pytest.mark.parametrize("with_profile", [False, True])
def test_inject_keeps_the_incident_wave(with_profile, rng):
    p = _kg(5, 4, 14, 0.004, src_k=5, k_beam=250.53)
    if with_profile:
        p["src_profile"] = rng.standard_normal((4, 5)) + 1j * rng.standard_normal((4, 5))
    t = 37 * p["dt"]
    state = _incident(p, t, p["src_profile"])
    assert _dev(fdtd3d.inject(state, t, p), state) < 1e-14
    planted = [_planted_plus_iw, _planted_flipped_k]
    if with_profile:
        planted.append(_planted_profile_dropped_from_pi)
    for bad in planted:
        assert _dev(bad(state, t, p), state) > 1e-3, bad.__name__


def test_inject_on_a_zero_state_writes_the_window(rng):
    # psi[k] = a_k psi_inc, pi[k] = -i w a_k psi_inc on K..K+2 (a = 0, 0.025, 0.05)
    p = _kg(4, 3, 12, 0.01, src_k=4)
    t = 3 * p["dt"]
    out = fdtd3d.inject(fdtd3d.zero_state(p), t, p)
    inc, pinc = _incident(p, t)
    w = _window(p)
    assert sorted(w) == [4, 5, 6] and w[4] == 0.0 and w[6] == pytest.approx(0.05, abs=1e-17)
    for k in range(p["nz"]):
        a = w.get(k, 0.0)
        assert np.allclose(out[0][k], a * inc[k], rtol=0, atol=1e-15)
        assert np.allclose(out[1][k], a * pinc[k], rtol=0, atol=1e-15 * p["e_sim"])


def _on_dispersion(p, k):
    """w of the positive-frequency z-uniform KG mode e^{i k z} on the discrete grid
    (V = 0): w^2 = c^2 (4/dz^2) sin^2(k dz/2) + m^2."""
    dz = (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0)
    return math.sqrt(C2 * 4.0 / dz ** 2 * math.sin(k * dz / 2) ** 2 + M_REST ** 2)


def _p4(x):
    return sum(x ** j / math.factorial(j) for j in range(5))


def test_incident_wave_on_the_dispersion_relation_steps_by_p4():
    # with w on the discrete dispersion relation the incident wave is an eigenmode of the
    # discrete KG operator, the source leaves it unchanged, and one RK4 step multiplies it
    # by P4(-i w dt) on every plane more than 4 cells from the frozen ends
    p = _kg(4, 4, 24, 0.004, src_k=8)
    p["k_beam"] = 150.0                       # k dz = 0.6
    p["e_sim"] = _on_dispersion(p, p["k_beam"])
    t = 11 * p["dt"]
    state = _incident(p, t)
    V = np.zeros(state[0].shape)
    got = fdtd3d.rk4_step(state, V, t, p)
    amp = _p4(-1j * p["e_sim"] * p["dt"])
    for g, s in zip(got, state):
        assert np.max(np.abs(g[5:-5] - amp * s[5:-5])) < 1e-12 * np.max(np.abs(s))
    # the injection with +i w leaves the eigenmode, and the step shows it
    bad = fdtd3d.rk4_step(_planted_plus_iw(state, t, p), V, t, dict(p, src_enabled=False))
    assert np.max(np.abs(bad[1][5:-5] - amp * state[1][5:-5])) > 1e-3 * np.max(np.abs(state[1]))


# ------------------------------------------------------------------ sponge
def _gamma_ref(p):
    """gamma_max r^p on the PML ramp, per plane with math (fdtd_3d.cu:96-106's ramp)."""
    out = []
    for z in _z(p):
        rb = max((p["zmin"] + p["pml_thick"]) - z, 0.0) / p["pml_thick"]
        rt = max(z - (p["zmax"] - p["pml_thick"]), 0.0) / p["pml_thick"]
        r = max(rb, rt)
        out.append(0.0 if r <= 0.0 else p["kg_gamma_max"] * r ** p["pml_power"])
    return np.array(out)


def test_sponge_profile_closed_form():
    p = _kg(4, 4, 41, 0.05, pml_thick=8 * 0.05, kg_gamma_max=3.0e5)
    g = fdtd3d.kg_gamma(p)
    assert g[0] == pytest.approx(3.0e5) and g[-1] == pytest.approx(3.0e5)
    assert g[4] == pytest.approx(3.0e5 * 0.5 ** 3)
    assert g[36] == pytest.approx(3.0e5 * 0.5 ** 3)
    assert np.all(g[8:33] < 1e-20)
    assert np.allclose(g, _gamma_ref(p), rtol=1e-13, atol=0)


def _lap_kg(p):
    """Periodic-x/y, 3-point-z Laplacian without stretch; rows of frozen planes empty."""
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    dx = (p["xmax"] - p["xmin"]) / (nx - 1.0)
    dy = (p["ymax"] - p["ymin"]) / (ny - 1.0)
    dz = (p["zmax"] - p["zmin"]) / (nz - 1.0)
    rows, cols, vals = [], [], []

    def ix(i, j, k):
        return (k * ny + j) * nx + i

    for k in range(1, nz - 1):
        for j in range(ny):
            for i in range(nx):
                r = ix(i, j, k)
                for c, v in ((ix((i + 1) % nx, j, k), 1 / dx ** 2),
                             (ix((i - 1) % nx, j, k), 1 / dx ** 2),
                             (ix(i, (j + 1) % ny, k), 1 / dy ** 2),
                             (ix(i, (j - 1) % ny, k), 1 / dy ** 2),
                             (ix(i, j, k + 1), 1 / dz ** 2), (ix(i, j, k - 1), 1 / dz ** 2),
                             (r, -2 / dx ** 2 - 2 / dy ** 2 - 2 / dz ** 2)):
                    rows.append(r)
                    cols.append(c)
                    vals.append(v)
    n = nx * ny * nz
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def _brute_force_step(state, V, t, p):
    """P4(dt M) b with M assembled from an independent gamma and b from an independent
    soft source (window, phase and -i w written out here)."""
    shape = (p["nz"], p["ny"], p["nx"])
    n = int(np.prod(shape))
    psi, pi = (s.copy() for s in state)
    inc, pinc = _incident(p, t, p.get("src_profile"))
    for k, a in _window(p).items():
        psi[k] = (1 - a) * psi[k] + a * inc[k]
        pi[k] = (1 - a) * pi[k] + a * pinc[k]
    mask = np.zeros(shape)
    mask[1:-1] = 1.0
    m = mask.ravel()
    g = (_gamma_ref(p)[:, None, None] * np.ones(shape)).ravel()
    v = V.ravel()
    top = sp.hstack([sp.csr_matrix((n, n)), sp.diags(m)])
    bot = sp.hstack([C2 * _lap_kg(p) - sp.diags((M_REST ** 2 - v ** 2) * m),
                     sp.diags((-2j * v - 2 * g) * m)])
    M = sp.vstack([top, bot]).tocsr()
    y = np.concatenate([psi.ravel(), pi.ravel()])
    out, term = y.copy(), y.copy()
    for j in range(1, 5):
        term = p["dt"] * (M @ term) / j
        out = out + term
    return out[:n].reshape(shape), out[n:].reshape(shape)


def test_step_with_sponge_and_source_equals_independent_p4(rng, monkeypatch):
    p = _kg(4, 5, 10, 0.01, pml_thick=0.035, src_k=3, kg_gamma_max=4.0e5)
    p["src_profile"] = rng.standard_normal((5, 4)) + 1j * rng.standard_normal((5, 4))
    shape = (10, 5, 4)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    pi = 7e5 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    V = 1e3 * rng.standard_normal(shape)
    t = 9 * p["dt"]
    assert np.count_nonzero(_gamma_ref(p)[1:-1]) >= 4      # the sponge reaches the interior
    want = _brute_force_step((psi, pi), V, t, p)
    assert _dev(fdtd3d.rk4_step((psi, pi), V, t, p), want) < 1e-12
    # planted: one extra power of the ramp in the oracle's sponge
    real = fdtd3d.kg_gamma
    monkeypatch.setattr(fdtd3d, "kg_gamma",
                        lambda q: real(dict(q, pml_power=q["pml_power"] + 1.0)))
    assert _dev(fdtd3d.rk4_step((psi, pi), V, t, p), want) > 1e-9


# ------------------------------------------------------------------ charge
def _charge(state, V):
    psi, pi = state
    return float(np.sum(np.imag(np.conj(psi) * pi) + V * np.abs(psi) ** 2))


def _charge_drift(p, state, V, T):
    """|Q(T) - Q(0)| / |Q(0)| after T / dt oracle steps."""
    n = int(round(T / p["dt"]))
    out = fdtd3d.run(state, V, n, p)
    q0 = _charge(state, V)
    return abs(_charge(out, V) - q0) / abs(q0)


def _closed_box_state(rng, shape):
    # a positive-frequency state (pi ~ -i m psi) plus an unaligned part, zero on the
    # frozen end planes so the box is closed
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    pi = -1j * M_REST * psi + 0.3 * M_REST * (rng.standard_normal(shape)
                                               + 1j * rng.standard_normal(shape))
    psi[[0, -1]] = 0
    pi[[0, -1]] = 0
    return psi, pi


def _conserves_charge(p, state, V, T):
    d1 = _charge_drift(p, state, V, T)
    d2 = _charge_drift(dict(p, dt=p["dt"] / 2), state, V, T)
    return d1 < 1e-5 and 20.0 < d1 / d2 < 45.0, (d1, d2)


def test_charge_is_conserved_to_rk4_order_without_sponge(rng, monkeypatch):
    p = _kg(6, 5, 12, 0.03, src_enabled=False, kg_gamma_max=0.0, dt=2e-7)
    shape = (12, 5, 6)
    state = _closed_box_state(rng, shape)
    V = 2e4 * rng.standard_normal(shape)
    T = 60 * p["dt"]
    ok, d = _conserves_charge(p, state, V, T)
    assert ok, d
    real = fdtd3d.rhs

    def planted(sign):
        def rhs(st, VV, q, cache=None):
            kpsi, kpi = real(st, VV, q, cache)
            kpi = kpi + (sign * 2j) * VV * st[1]   # -2iV pi -> 0 (sign 1) or +2iV pi (2)
            kpi[[0, -1]] = 0
            return kpsi, kpi
        return rhs

    for sign in (1.0, 2.0):
        monkeypatch.setattr(fdtd3d, "rhs", planted(sign))
        ok, d = _conserves_charge(p, state, V, T)
        assert not ok, (sign, d)
    monkeypatch.setattr(fdtd3d, "rhs", real)


def test_sponge_drains_the_charge_at_the_rate_of_the_damping_term(rng, monkeypatch):
    # dQ/dt = -2 sum gamma Im(conj(psi) pi): for a positive-frequency field in the sponge
    # Q (= -(w - V) |psi|^2 < 0) rises towards 0; one short step moves it by dt dQ/dt
    p = _kg(4, 4, 16, 0.03, src_enabled=False, pml_thick=0.2, kg_gamma_max=2.0e5, dt=2e-9)
    shape = (16, 4, 4)
    state = _closed_box_state(rng, shape)
    V = 1e3 * rng.standard_normal(shape)
    g = _gamma_ref(p)[:, None, None]

    def rate_ratio():
        q0 = _charge(state, V)
        q1 = _charge(fdtd3d.rk4_step(state, V, 0.0, p), V)
        qdot = -2.0 * float(np.sum(g * np.imag(np.conj(state[0]) * state[1])))
        return (q1 - q0) / (p["dt"] * qdot), q0, q1

    r, q0, q1 = rate_ratio()
    assert q0 < 0 and abs(q1) < abs(q0)
    assert r == pytest.approx(1.0, abs=2e-2)
    real = fdtd3d.kg_gamma
    for scale in (0.5, -1.0):     # gamma pi for 2 gamma pi; a growing sponge
        monkeypatch.setattr(fdtd3d, "kg_gamma", lambda q, s=scale: s * real(q))
        assert rate_ratio()[0] != pytest.approx(1.0, abs=2e-2), scale
    monkeypatch.setattr(fdtd3d, "kg_gamma", real)
    # and with the sponge on for a while, the charge keeps draining
    out = fdtd3d.run(state, V, 200, dict(p, dt=2e-8))
    assert abs(_charge(out, V)) < abs(q0)


# ------------------------------------------------------- forward propagation
def _phase_slope(field_z, z):
    ph = np.unwrap(np.angle(field_z))
    slope, icpt = np.polyfit(z, ph, 1)
    res = ph - (slope * z + icpt)
    r2 = 1.0 - np.sum(res ** 2) / np.sum((ph - ph.mean()) ** 2)
    return slope, r2


def _forward_case(inject_fn=None, monkeypatch=None):
    lam = 0.1
    k = 2 * math.pi / lam
    e_tot = math.sqrt((1973.2698 * k) ** 2 + M_REST ** 2)
    v_g = C2 * k / e_tot
    T = 0.048
    gmax = math.log(1e6) * 4.0 * v_g / T
    p = _kg(4, 4, 80, 0.008, pml_thick=T, src_k=10, k_beam=k, e_sim=e_tot,
            kg_gamma_max=gmax, dt=1e-6)
    if inject_fn is not None:
        monkeypatch.setattr(fdtd3d, "inject", inject_fn)
    n = 1400
    out = fdtd3d.run(fdtd3d.zero_state(p), np.zeros((80, 4, 4)), n, p)
    t = n * p["dt"]
    z = np.array(_z(p))
    demod = out[0][:, 2, 2] * np.exp(1j * e_tot * t)
    sel = slice(20, 60)
    return k, z[sel], demod[sel]


def _rk4_wavenumber(w, dt, dz):
    """Wavenumber of the wave a drive e^{-i w t} launches on the fully discrete grid:
    the mode frequency W whose RK4 step turns the phase by w dt (arg P4(-i W dt) = -w dt),
    then c^2 (4/dz^2) sin^2(k_d dz/2) = W^2 - m^2."""
    lo, hi = w * 0.99, w * 1.05
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if -np.angle(_p4(-1j * mid * dt)) < w * dt:
            lo = mid
        else:
            hi = mid
    W = 0.5 * (lo + hi)
    K = math.sqrt((W * W - M_REST ** 2) / C2)
    return 2.0 / dz * math.asin(K * dz / 2.0)


def test_kg_source_propagates_forward(monkeypatch):
    # tests/test_fdtd_3d.py:312-323 of the reference: the demodulated phase of the
    # source's wave rises along +z; here also with the fully discrete wavenumber
    # (k dz = 0.5, w dt = 0.53: k_d = 1.0215 k) and a flat amplitude
    k, z, f = _forward_case()
    slope, r2 = _phase_slope(f, z)
    amp = np.abs(f)
    assert slope > 0 and r2 > 0.999
    e_tot = math.sqrt(C2 * k * k + M_REST ** 2)
    assert slope == pytest.approx(_rk4_wavenumber(e_tot, 1e-6, 0.008), rel=3e-3)
    assert amp.min() > 0.5 * amp.max() and amp.max() > 0.1
    # planted: +i w in the pi injection drives the backward branch off resonance
    real = fdtd3d.inject
    _, zb, fb = _forward_case(lambda st, t, p: _planted_plus_iw_with(real, st, t, p),
                              monkeypatch)
    assert np.abs(fb).max() < 0.1 * amp.max()


def _planted_plus_iw_with(inject, state, t, p):
    psi, pi = inject(state, t, p)
    inc, _ = _incident(p, t, p.get("src_profile"))
    pi = pi.copy()
    for k, a in _window(p).items():
        pi[k] += a * 2j * p["e_sim"] * inc[k]
    return psi, pi


# ------------------------------------------------------------ float32 clock
def _round_f32(x):
    """Nearest binary32 value of the rational x (ties to even); normal range."""
    x = Fraction(x)
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    # a in [2^e, 2^(e+1)): 24 significant bits -> ulp 2^(e-23)
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(a, ulp)
    if r > ulp / 2 or (r == ulp / 2 and q % 2 == 1):
        q += 1
    return s * q * ulp


def test_round_f32_matches_known_binary32_values():
    assert _round_f32(Fraction(1, 10)) == Fraction(13421773, 2 ** 27)   # 0x3DCCCCCD
    assert _round_f32(16777217) == 16777216                            # tie to even
    assert _round_f32(16777219) == 16777220                            # tie to even
    assert _round_f32(Fraction(5, 10 ** 7)) == Fraction(np.float32(5e-7).item())


@pytest.mark.parametrize("dt", [5e-7, 2e-6, 1e-6, 3.7e-7])
def test_ref_time_is_the_float32_product(dt):
    # fdtd_3d.cu:365: float time = (frame * STEPS_PER_FRAME + s) * DT, an int times a float
    dt32 = _round_f32(Fraction(dt))
    ns = [0, 1, 2, 3, 7, 100, 999, 1000, 12345, 19999, 20000, 2 ** 24 - 1, 2 ** 24 + 1,
          2 ** 24 + 3, 33554431, 123456789]
    mism = 0
    for n in ns:
        want = float(_round_f32(_round_f32(n) * dt32))
        assert fdtd3d.ref_time(n, dt) == want, n
        # a double product rounded once (the plausible mistake) differs for some n
        if float(np.float32(n * dt)) != want:
            mism += 1
    assert mism > 0
