"""ctypes front of ``oracle/fdtd3d_omp.c``: the FP64 C + OpenMP restatement of
``oracle.fdtd3d``'s RK4 step (the same reference lines, ``fdtd_3d.cu:31-273,
363-382``).

TEST INFRASTRUCTURE ONLY: bench.py's multi-threaded CPU baseline, checked against the
NumPy oracle in ``tests/test_oracle_c.py``.  The per-plane coefficients (L_z weights,
sponge, source window) come from ``oracle.fdtd3d`` itself; the C code runs the
stencil sweeps.
"""
import ctypes
import os
import subprocess

import numpy as np

from . import fdtd3d

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(_HERE, "fdtd3d_omp.c")
LIB = os.path.join(_HERE, "libfdtd3d_omp.so")
# -fcx-limited-range: textbook complex products (C's Annex G inf/nan handling calls a
# library routine per product); portable x86-64 code, the library runs on other hosts
CFLAGS = ["-O3", "-std=gnu11", "-fopenmp", "-fcx-limited-range", "-fPIC", "-shared"]

_lib = None


def build(force=False):
    """Compile the library in place (gcc) when missing or older than its source."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        tmp = LIB + ".tmp"
        subprocess.run(["gcc", *CFLAGS, "-o", tmp, SRC], check=True)
        os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(build())
        c, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_rk4_run.argtypes = [i, i, i, i, d, d, c, c, c, d, d, d, d, i, c, i, c, c, c, c,
                                     d, c, i, c, c]
        L.oracle_rk4_run.restype = i
        L.oracle_set_threads.argtypes = [i]
        L.oracle_max_threads.restype = i
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def run(state, V, nsteps, p, t0=0.0, step0=0, f32_time=False, threads=0):
    """As ``oracle.fdtd3d.run`` (without the callback): a new state after nsteps steps.
    threads > 0 sets the OpenMP thread count."""
    L = lib()
    kg = p["equation"] != "schrodinger"
    nz, ny, nx = p["nz"], p["ny"], p["nx"]
    dx, dy, _ = fdtd3d.spacing(p)
    cache = fdtd3d._cache(p)
    cp = np.ascontiguousarray(cache["cp"], dtype=np.complex128)
    cm = np.ascontiguousarray(cache["cm"], dtype=np.complex128)
    gam = np.ascontiguousarray(cache.get("gamma", np.zeros(nz)), dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    assert V.shape == (nz, ny, nx)
    if kg:
        psi = np.array(state[0], dtype=np.complex128, order="C")
        pi = np.array(state[1], dtype=np.complex128, order="C")
    else:
        psi = np.array(state, dtype=np.complex128, order="C")
        pi = None
    assert psi.shape == (nz, ny, nx) and (pi is None or pi.shape == psi.shape)
    # the source: window planes and weights, e^{i k (z_s - z_K)}, e^{-i E t_n} per step
    if p.get("src_enabled", True):
        ks, ws = fdtd3d.source_weights(p)
    else:
        ks, ws = [], []
    z = fdtd3d.z_coords(p)
    sk = int(p["src_k"])
    zk = z[sk] if 0 <= sk < nz else p["zmin"] + sk * (p["zmax"] - p["zmin"]) / (nz - 1.0)
    src_k = np.array(ks, dtype=np.int32)
    src_w = np.array(ws, dtype=np.float64)
    src_ph = np.exp(1j * p["k_beam"] * (z[ks] - zk)).astype(np.complex128) if ks else \
        np.zeros(1, dtype=np.complex128)
    n = np.arange(step0, step0 + nsteps)
    if f32_time:
        t = np.array([fdtd3d.ref_time(int(m), p["dt"]) for m in n])
    else:
        t = t0 + n * p["dt"]
    step_ph = np.exp(-1j * p["e_sim"] * t).astype(np.complex128)
    prof = p.get("src_profile")
    prof = None if prof is None else np.ascontiguousarray(prof, dtype=np.complex128)
    L.oracle_set_threads(int(threads))
    rc = L.oracle_rk4_run(
        1 if kg else 0, nx, ny, nz, 1.0 / (dx * dx), 1.0 / (dy * dy) if ny > 1 else 0.0,
        _p(cp), _p(cm), _p(gam), float(p.get("A", 0.0)),
        float(p.get("hbar_c", 0.0)) ** 2 if kg else 0.0,
        float(p.get("rest_energy", 0.0)) ** 2 if kg else 0.0, float(p["dt"]),
        1 if p.get("end_planes", "frozen") == "zero" else 0, _p(V), len(ks), _p(src_k),
        _p(src_w), _p(src_ph), _p(prof), float(p["e_sim"]), _p(step_ph), int(nsteps),
        _p(psi), _p(pi))
    if rc != 0:
        raise MemoryError("oracle_rk4_run: scratch allocation failed")
    return (psi, pi) if kg else psi


def max_threads():
    return int(lib().oracle_max_threads())
