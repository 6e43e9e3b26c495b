"""ctypes binding of ``libwapfdtd.so`` (the C ABI declared in include/wap_fdtd.h).

torch is imported first so that the HIP runtime it ships is the one the
library binds to (both carry the soname libamdhip64.so.7): device pointers
and streams are then shared between torch and the solver.  There is no CPU
fallback: if the library is missing or cannot be loaded, ``lib()`` raises.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime before the solver library)

from .build import LIB_PATH

WAP_OK = 0
EQ = {"schrodinger": 0, "klein_gordon": 1}
PREC = {"f64": 0, "f32": 1, "f32_kahan": 2}
TIME = {"exact": 0, "f32": 1}
SCHED = {"auto": 0, "stages": 1, "kg_2pass": 2, "kg_fused": 3, "schr_fused": 4, "schr_1pass": 5}
DT_F32, DT_F64, DT_C64, DT_C128 = 0, 1, 2, 3
WAP_STREAM_ORDERED = 2   # on_device value: device memory, copied on the handle's stream

_c_int = ctypes.c_int
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class WapDesc(ctypes.Structure):
    """Mirror of ``wap_desc`` (include/wap_fdtd.h)."""
    _fields_ = [
        ("nx", _c_i32), ("ny", _c_i32), ("nz", _c_i32),
        ("equation", _c_i32), ("precision", _c_i32), ("time_mode", _c_i32),
        ("src_enabled", _c_i32), ("src_k", _c_i32), ("src_band", _c_i32),
        ("tile_y", _c_i32), ("tile_z", _c_i32), ("schedule", _c_i32),
        ("end_planes", _c_i32), ("reserved0", _c_i32),
        ("xmin", _c_dbl), ("xmax", _c_dbl), ("ymin", _c_dbl), ("ymax", _c_dbl),
        ("zmin", _c_dbl), ("zmax", _c_dbl),
        ("dt", _c_dbl), ("lap_coeff", _c_dbl), ("e_sim", _c_dbl), ("k_beam", _c_dbl),
        ("hbar_c", _c_dbl), ("rest_energy", _c_dbl),
        ("pml_thick", _c_dbl), ("pml_sigma0", _c_dbl), ("pml_theta", _c_dbl),
        ("pml_power", _c_dbl), ("kg_gamma_max", _c_dbl), ("src_alpha", _c_dbl),
    ]


# name -> (restype, argtypes); every symbol include/wap_fdtd.h declares
SIGNATURES = {
    "wap_abi_version": (_c_int, []),
    "wap_last_error": (ctypes.c_char_p, []),
    "wap_create": (_c_int, [ctypes.POINTER(WapDesc), ctypes.POINTER(_vp)]),
    "wap_destroy": (_c_int, [_vp]),
    "wap_set_stream": (_c_int, [_vp, _vp]),
    "wap_set_potential": (_c_int, [_vp, _vp, _c_int, _c_int]),
    "wap_set_state": (_c_int, [_vp, _vp, _vp, _c_int, _c_int]),
    "wap_get_state": (_c_int, [_vp, _vp, _vp, _c_int, _c_int]),
    "wap_set_source_profile": (_c_int, [_vp, _vp, _c_int, _c_int]),
    "wap_step": (_c_int, [_vp, _c_i64, _c_i64, _c_dbl]),
    "wap_step_profiled": (_c_int, [_vp, _c_i64, _c_i64, _c_dbl, ctypes.POINTER(_c_dbl),
                                   ctypes.POINTER(_c_dbl)]),
    "wap_get_plane": (_c_int, [_vp, _c_int, _vp, _c_int, _c_int]),
    "wap_accumulate_plane": (_c_int, [_vp, _c_int, _c_dbl, _c_dbl, _vp]),
    "wap_norm2": (_c_int, [_vp, ctypes.POINTER(_c_dbl)]),
    "wap_synchronize": (_c_int, [_vp]),
    "wap_device_bytes": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "wap_launch_shape": (_c_int, [_vp, ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i32),
                                  ctypes.POINTER(_c_i64)]),
    "wap_batch_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, ctypes.POINTER(_vp)]),
    "wap_batch_destroy": (_c_int, [_vp]),
    "wap_batch_set_stream": (_c_int, [_vp, _vp]),
    "wap_batch_tiles": (_c_int, [_vp, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i32)]),
    "wap_batch_step": (_c_int, [_vp, _c_i64, _c_i64]),
    "wap_batch_capture": (_c_int, [_vp, _c_i64]),
    "wap_batch_replay": (_c_int, [_vp, _c_i64, ctypes.POINTER(_c_dbl)]),
    "wap_batch_synchronize": (_c_int, [_vp]),
    "wap_batch_set_create": (_c_int, [ctypes.POINTER(_vp), _c_int, ctypes.POINTER(_vp)]),
    "wap_batch_set_destroy": (_c_int, [_vp]),
    "wap_batch_set_capture": (_c_int, [_vp, _c_i64]),
    "wap_batch_set_replay": (_c_int, [_vp, _c_i64, ctypes.POINTER(_c_dbl)]),
    "wap_batch_set_synchronize": (_c_int, [_vp]),
    "wap_adjoint_create": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp)]),
    "wap_adjoint_workspace_bytes": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_i64)]),
    "wap_adjoint_create_in": (_c_int, [_vp, _c_int, _vp, _c_i64, ctypes.POINTER(_vp)]),
    "wap_adjoint_destroy": (_c_int, [_vp]),
    "wap_adjoint_reset": (_c_int, [_vp]),
    "wap_adjoint_record": (_c_int, [_vp, _c_i64, _c_dbl, _c_int]),
    "wap_adjoint_step": (_c_int, [_vp, _c_int]),
    "wap_adjoint_seed": (_c_int, [_vp, _c_int, _vp, _c_dbl, _c_dbl]),
    "wap_adjoint_get_grad": (_c_int, [_vp, _vp, _c_int]),
    "wap_stream_bandwidth": (_c_int, [_c_i64, _c_int, ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]),
    "wap_slice_transmit": (_c_int, [_vp, _c_int, _vp, _c_int, _c_i64, _c_dbl, _c_dbl, _c_dbl,
                                    _c_dbl, _vp]),
    "wap_atom_potential": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_dbl, _c_dbl, _vp,
                                    _c_int, _vp]),
    "wap_atom_potential_static": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_dbl,
                                           _c_dbl, _vp, _c_int, _vp]),
}

_LIB = None


class WapError(RuntimeError):
    """A C-ABI call returned a negative status."""


def lib():
    """Load (once) and return the solver library; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m wideanglepropagation_amd.build)")
    handle = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = handle
    return _LIB


def check(rc, what=""):
    if rc != WAP_OK:
        msg = lib().wap_last_error().decode(errors="replace")
        raise WapError(f"{what} failed ({rc}): {msg}")
    return rc


def make_desc(p):
    """FDTDParams -> WapDesc."""
    d = WapDesc()
    d.nx, d.ny, d.nz = int(p.nx), int(p.ny), int(p.nz)
    d.equation = EQ[p.equation]
    d.precision = PREC[p.precision]
    d.time_mode = TIME[p.time_mode]
    d.src_enabled = 1 if p.src_enabled else 0
    d.src_k = int(p.src_k)
    d.src_band = int(p.src_band)
    d.tile_y = int(p.tile_y)
    d.tile_z = int(p.tile_z)
    d.schedule = SCHED[p.schedule]
    d.end_planes = {"frozen": 0, "zero": 1}[p.end_planes]
    (d.xmin, d.xmax, d.ymin, d.ymax, d.zmin, d.zmax) = [float(v) for v in p.extent]
    d.dt = float(p.dt)
    d.lap_coeff = float(p.lap_coeff)
    d.e_sim = float(p.e_sim)
    d.k_beam = float(p.k_beam)
    d.hbar_c = float(p.hbar_c)
    d.rest_energy = float(p.rest_energy)
    d.pml_thick = float(p.pml_thick)
    d.pml_sigma0 = float(p.pml_sigma0)
    d.pml_theta = float(p.pml_theta)
    d.pml_power = float(p.pml_power)
    d.kg_gamma_max = float(p.kg_gamma_max)
    d.src_alpha = float(p.src_alpha)
    return d


def stream_bandwidth(nbytes=2 << 30, reps=10):
    """(copy GB/s, read GB/s) of STREAM kernels on nbytes-sized buffers."""
    c, r = ctypes.c_double(), ctypes.c_double()
    check(lib().wap_stream_bandwidth(int(nbytes), int(reps), ctypes.byref(c), ctypes.byref(r)),
          "wap_stream_bandwidth")
    return c.value, r.value
