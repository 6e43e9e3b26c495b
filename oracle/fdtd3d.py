"""FP64 NumPy restatement of the reference 3D RK4 FDTD stepper (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``wide_angle_propagation/fdtd_3d.cu`` of the reference:
  * grid / index / spacing        ``fdtd_3d.cu:31-33, 81-83, 98``
  * PML sigma(z), a(z) = 1/S_z     ``fdtd_3d.cu:96-123``
  * soft plane-wave source         ``fdtd_3d.cu:126-161``
  * RK4 stage (stencil, k = -iH y) ``fdtd_3d.cu:164-257``
  * final update                  ``fdtd_3d.cu:259-273``
  * per-step order and time       ``fdtd_3d.cu:363-382``
with the stage-2/3 in-place race of ``fdtd_3d.cu:374-377`` removed: every
stage reads a stage input that no thread of the same stage writes, i.e. the
intended classical RK4.  Planes k = 0 and nz-1 are frozen at the (injected)
base value, as ``fdtd_3d.cu:184-190`` and ``:268`` make them.

2D (``fdtd.cu``, the reference's ``RelativisticSolver``) is the same
stepper on a grid with ny = 1 (no y term; the reference's y is z here) and
``end_planes = "zero"``: the end rows of the stage outputs are 0
(``fdtd.cu:189-198``) instead of the base value, while the state's end rows
still keep b (``fdtd.cu:272``).  Atomic units: A = 0.5 (``fdtd.cu:250-252``).

Klein-Gordon (not in the reference; the build's extension): the first-order
system in (psi, pi = d psi / dt), hbar = 1, time in hbar/eV:

    d psi/dt = pi
    d pi /dt = (hbar c)^2 Lap psi - (m^2 - V^2) psi - 2 i V pi - 2 gamma(z) pi

Lap is the same periodic-x/y, 3-point-z operator without coordinate
stretching; the absorbing layer is the damping rate gamma(z) = gamma_max r^p
on the same ramp r(z) as the reference PML (a complex coordinate stretch
makes the negative-frequency KG branch grow, so KG uses this sponge).  The
soft source injects psi_inc = P(x,y) e^{i(k (z - z_K) - w t)} into psi and
-i w psi_inc into pi with the same window.

Parameters are a plain dict (``p``) with keys:
  nx ny nz, xmin xmax ymin ymax zmin zmax, dt, A, e_sim, k_beam,
  equation ("schrodinger" | "klein_gordon"), hbar_c, rest_energy,
  pml_thick pml_sigma0 pml_theta pml_power, kg_gamma_max,
  src_enabled, src_k, src_band, src_alpha, src_profile (None or (ny,nx)).
"""
import numpy as np


def spacing(p):
    """``fdtd_3d.cu:31-33``: N-1 intervals on every axis (also periodic x, y).
    ny = 1 is the 2D grid of ``fdtd.cu`` (no y term)."""
    dx = (p["xmax"] - p["xmin"]) / (p["nx"] - 1.0)
    dy = (p["ymax"] - p["ymin"]) / (p["ny"] - 1.0) if p["ny"] > 1 else np.inf
    dz = (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0)
    return dx, dy, dz


def z_coords(p):
    """``fdtd_3d.cu:98, 141-142``."""
    k = np.arange(p["nz"], dtype=np.float64)
    return p["zmin"] + k * (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0)


def pml_ramp(p):
    """r(z) of ``fdtd_3d.cu:100-102``: 0 inside, rising to 1 at both z faces."""
    z = z_coords(p)
    t = p["pml_thick"]
    rb = np.maximum((p["zmin"] + t) - z, 0.0) / t
    rt = np.maximum(z - (p["zmax"] - t), 0.0) / t
    return np.maximum(rb, rt)


def pml_sigma(p):
    """``fdtd_3d.cu:96-106``: sigma0 r^p where r > 0, else 0."""
    r = pml_ramp(p)
    sig = np.zeros_like(r)
    m = r > 0.0
    sig[m] = p["pml_sigma0"] * r[m] ** p["pml_power"]
    return sig


def pml_a(p):
    """``fdtd_3d.cu:109-123``: a = 1/S, S = 1 + sigma e^{i theta} (a = 1 where sigma = 0)."""
    sig = pml_sigma(p)
    a = np.ones(p["nz"], dtype=np.complex128)
    m = sig > 0.0
    a[m] = 1.0 / (1.0 + sig[m] * np.exp(1j * p["pml_theta"]))
    return a


def kg_gamma(p):
    """Sponge damping rate of the KG extension on the PML ramp."""
    r = pml_ramp(p)
    g = np.zeros_like(r)
    m = r > 0.0
    g[m] = p["kg_gamma_max"] * r[m] ** p["pml_power"]
    return g


def z_coefficients(p):
    """Per-plane complex weights of L_z (``fdtd_3d.cu:215-234``).

    L_z y_k = cp_k (y_{k+1} - y_k) - cm_k (y_k - y_{k-1}) with
    cp_k = a_k (a_k + a_{k+1})/2 / dz^2 and cm_k = a_k (a_{k-1} + a_k)/2 / dz^2.
    Values on k = 0 and nz-1 are unused (frozen planes).
    """
    _, _, dz = spacing(p)
    if p["equation"] == "schrodinger":
        a = pml_a(p)
    else:
        a = np.ones(p["nz"], dtype=np.complex128)
    cp = np.zeros(p["nz"], dtype=np.complex128)
    cm = np.zeros(p["nz"], dtype=np.complex128)
    cp[1:-1] = a[1:-1] * 0.5 * (a[1:-1] + a[2:]) / (dz * dz)
    cm[1:-1] = a[1:-1] * 0.5 * (a[:-2] + a[1:-1]) / (dz * dz)
    return cp, cm


def laplacian(y, p, cp=None, cm=None):
    """Lx + Ly + Lz on interior planes; zero on k = 0, nz-1 (``fdtd_3d.cu:192-238``)."""
    dx, dy, _ = spacing(p)
    if cp is None:
        cp, cm = z_coefficients(p)
    lap = np.zeros_like(y)
    yc = y[1:-1]
    lap[1:-1] = ((np.roll(yc, -1, axis=2) - yc) + (np.roll(yc, 1, axis=2) - yc)) / (dx * dx)
    if p["ny"] > 1:
        lap[1:-1] += ((np.roll(yc, -1, axis=1) - yc) + (np.roll(yc, 1, axis=1) - yc)) / (dy * dy)
    lap[1:-1] += (cp[1:-1, None, None] * (y[2:] - yc)
                  - cm[1:-1, None, None] * (yc - y[:-2]))
    return lap


def rhs(state, V, p, cache=None):
    """k = F(y): Schrodinger k = -i H y, H = -A Lap + V (``fdtd_3d.cu:242-248``);
    KG as in the module docstring.  Zero on the frozen planes."""
    if cache is None:
        cache = _cache(p)
    cp, cm = cache["cp"], cache["cm"]
    if p["equation"] == "schrodinger":
        y = state
        h = -p["A"] * laplacian(y, p, cp, cm) + V * y
        k = -1j * h
        k[0] = 0.0
        k[-1] = 0.0
        return k
    psi, pi = state
    c2 = p["hbar_c"] ** 2
    m2 = p["rest_energy"] ** 2
    g = cache["gamma"][:, None, None]
    kpsi = pi.copy()
    kpi = c2 * laplacian(psi, p, cp, cm) - (m2 - V * V) * psi - 2j * V * pi - 2.0 * g * pi
    for arr in (kpsi, kpi):
        arr[0] = 0.0
        arr[-1] = 0.0
    return (kpsi, kpi)


def _cache(p):
    cp, cm = z_coefficients(p)
    c = {"cp": cp, "cm": cm}
    if p["equation"] != "schrodinger":
        c["gamma"] = kg_gamma(p)
    return c


def source_weights(p):
    """Window weights a_k on planes K..K+band-1 (``fdtd_3d.cu:134, 150-156``)."""
    band = int(p["src_band"])
    ks = []
    ws = []
    for k in range(int(p["src_k"]), int(p["src_k"]) + band):
        if k < 0 or k >= p["nz"]:
            continue
        s = (k - p["src_k"]) / (band - 1.0) if band > 1 else 0.0
        win = 0.5 * (1.0 - np.cos(np.pi * s))
        ks.append(k)
        ws.append(p["src_alpha"] * win)
    return ks, ws


def inject(state, t, p):
    """Soft source at time t (``fdtd_3d.cu:126-161``); returns the injected state."""
    if not p.get("src_enabled", True):
        return state
    z = z_coords(p)
    zk = z[int(p["src_k"])] if 0 <= int(p["src_k"]) < p["nz"] else (
        p["zmin"] + p["src_k"] * (p["zmax"] - p["zmin"]) / (p["nz"] - 1.0))
    prof = p.get("src_profile")
    ks, ws = source_weights(p)
    if p["equation"] == "schrodinger":
        psi = state.copy()
    else:
        psi, pi = state[0].copy(), state[1].copy()
    for k, a in zip(ks, ws):
        phase = p["k_beam"] * (z[k] - zk) - p["e_sim"] * t
        inc = np.exp(1j * phase)
        if prof is not None:
            inc = inc * prof
        psi[k] = (1.0 - a) * psi[k] + a * inc
        if p["equation"] != "schrodinger":
            pi[k] = (1.0 - a) * pi[k] + a * (-1j * p["e_sim"]) * inc
    if p["equation"] == "schrodinger":
        return psi
    return (psi, pi)


def _axpy(base, c, k, zero_ends=False):
    """A stage input b + c k.  zero_ends: the 2D reference writes 0 into the
    stage outputs on its end rows (``fdtd.cu:189-198``)."""
    if isinstance(base, tuple):
        return tuple(_axpy(b, c, kk, zero_ends) for b, kk in zip(base, k))
    y = base + c * k
    if zero_ends:
        y[0] = 0.0
        y[-1] = 0.0
    return y


def rk4_step(state, V, t, p, cache=None):
    """One step: inject at t, then classical RK4 (``fdtd_3d.cu:363-382``).

    psi_{n+1} = b + dt/6 (k1 + 2 k2 + 2 k3 + k4), b = injected psi_n; the
    frozen planes keep b.
    """
    if cache is None:
        cache = _cache(p)
    dt = p["dt"]
    b = inject(state, t, p)
    ze = p.get("end_planes", "frozen") == "zero"
    k1 = rhs(b, V, p, cache)
    k2 = rhs(_axpy(b, 0.5 * dt, k1, ze), V, p, cache)
    k3 = rhs(_axpy(b, 0.5 * dt, k2, ze), V, p, cache)
    k4 = rhs(_axpy(b, dt, k3, ze), V, p, cache)
    if isinstance(b, tuple):
        return tuple(bb + (dt / 6.0) * (a1 + 2.0 * a2 + 2.0 * a3 + a4)
                     for bb, a1, a2, a3, a4 in zip(b, k1, k2, k3, k4))
    return b + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def ref_time(n, dt):
    """``fdtd_3d.cu:365``: float time = (frame*SPF + s) * DT evaluated in float32."""
    return float(np.float32(np.float32(n) * np.float32(dt)))


def run(state, V, nsteps, p, t0=0.0, step0=0, callback=None, f32_time=False):
    """Advance ``nsteps`` steps; step n uses t = t0 + (step0 + n) dt
    (or the reference's float32 clock when ``f32_time``).

    ``callback(n_done, state)`` is called after each step when given.
    """
    cache = _cache(p)
    V = np.asarray(V, dtype=np.float64)
    for n in range(nsteps):
        if f32_time:
            t = ref_time(step0 + n, p["dt"])
        else:
            t = t0 + (step0 + n) * p["dt"]
        state = rk4_step(state, V, t, p, cache)
        if callback is not None:
            callback(n + 1, state)
    return state


def zero_state(p):
    shape = (p["nz"], p["ny"], p["nx"])
    if p["equation"] == "schrodinger":
        return np.zeros(shape, dtype=np.complex128)
    return (np.zeros(shape, dtype=np.complex128), np.zeros(shape, dtype=np.complex128))


def norm2(state):
    """sum |psi|^2 over the grid (FP64)."""
    psi = state[0] if isinstance(state, tuple) else state
    return float(np.sum(np.abs(psi) ** 2))
