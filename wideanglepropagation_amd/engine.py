"""Python handle over one device-resident FDTD grid (wraps the C ABI).

``FDTD3D`` owns a ``wap_handle``: the state, the RK4 stage buffers, V and the
coefficient tables live in HBM for the lifetime of the object; ``step`` only
launches kernels.  Host arrays (numpy) and device tensors (torch, on the
current HIP device) are both accepted wherever a field is passed in.  Work is
ordered on torch's current stream at construction time, so torch operations
on tensors the solver fills (exit-plane accumulators, far fields) need no
extra synchronisation.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, make_desc, DT_F32, DT_F64, DT_C64, DT_C128, WAP_STREAM_ORDERED


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP FDTD solver needs a visible AMD GPU (torch.cuda is unavailable)")


def _ptr_of(x, kind, expect_numel):
    """Return (keepalive, pointer, dtype code, on_device) for a numpy array or torch tensor."""
    if isinstance(x, torch.Tensor):
        t = x.contiguous()
        if kind == "real":
            if t.dtype not in (torch.float32, torch.float64):
                t = t.to(torch.float64)
            code = DT_F32 if t.dtype == torch.float32 else DT_F64
        else:
            if t.dtype not in (torch.complex64, torch.complex128):
                t = t.to(torch.complex128)
            code = DT_C64 if t.dtype == torch.complex64 else DT_C128
        if t.numel() != expect_numel:
            raise ValueError(f"expected {expect_numel} elements, got {t.numel()}")
        on_dev = 1 if t.is_cuda else 0
        return t, ctypes.c_void_p(t.data_ptr()), code, on_dev
    a = np.asarray(x)
    if kind == "real":
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float64)
        code = DT_F32 if a.dtype == np.float32 else DT_F64
    else:
        if a.dtype not in (np.complex64, np.complex128):
            a = a.astype(np.complex128)
        code = DT_C64 if a.dtype == np.complex64 else DT_C128
    a = np.ascontiguousarray(a)
    if a.size != expect_numel:
        raise ValueError(f"expected {expect_numel} elements, got {a.size}")
    return a, ctypes.c_void_p(a.ctypes.data), code, 0


class FDTD3D:
    """One device-resident grid of the RK4 FDTD stepper."""

    def __init__(self, params, V=None, psi0=None, pi0=None, source_profile=None, t0=0.0):
        if params.time_mode == "f32" and float(t0) != 0.0:
            # the reference's float32 clock (fdtd_3d.cu:365) has no time origin
            raise ValueError("t0 must be 0 with time_mode='f32'")
        _require_gpu()
        self.params = params
        self._L = lib()
        self._desc = make_desc(params)
        h = ctypes.c_void_p()
        check(self._L.wap_create(ctypes.byref(self._desc), ctypes.byref(h)), "wap_create")
        self._h = h
        self.stream = torch.cuda.current_stream()
        check(self._L.wap_set_stream(self._h, ctypes.c_void_p(self.stream.cuda_stream)),
              "wap_set_stream")
        self.step_index = 0
        self.t0 = float(t0)
        if V is not None:
            self.set_potential(V)
        if psi0 is not None:
            self.set_state(psi0, pi0)
        if source_profile is not None:
            self.set_source_profile(source_profile)

    # ----------------------------------------------------------------- setup
    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("FDTD3D handle is closed")
        return self._h

    def set_potential(self, V):
        keep, ptr, code, dev = _ptr_of(V, "real", self.params.voxels)
        check(self._L.wap_set_potential(self.handle, ptr, code, dev), "wap_set_potential")
        del keep

    def set_state(self, psi, pi=None, stream_ordered=False):
        """stream_ordered: device tensors are copied on the solver's stream (torch's stream
        when the solver was made) without waiting for the copy."""
        n = self.params.voxels
        if psi is None:
            check(self._L.wap_set_state(self.handle, None, None, DT_C128, 0), "wap_set_state")
            return
        kp, pp, code, dev = _ptr_of(psi, "complex", n)
        if pi is not None:
            if isinstance(pi, torch.Tensor):
                pi = pi.to(kp.dtype) if isinstance(kp, torch.Tensor) else pi.cpu().numpy()
            kq, pq, code_q, dev_q = _ptr_of(pi, "complex", n)
            if (code_q, dev_q) != (code, dev):
                raise ValueError("psi and pi must have the same dtype and device")
        else:
            kq, pq = None, None
        if stream_ordered and dev:
            dev = WAP_STREAM_ORDERED
        check(self._L.wap_set_state(self.handle, pp, pq, code, dev), "wap_set_state")
        del kp, kq

    def set_source_profile(self, prof):
        if prof is None:
            check(self._L.wap_set_source_profile(self.handle, None, DT_C128, 0), "profile")
            return
        keep, ptr, code, dev = _ptr_of(prof, "complex", self.params.nx * self.params.ny)
        check(self._L.wap_set_source_profile(self.handle, ptr, code, dev), "profile")
        del keep

    # ------------------------------------------------------------------ run
    def step(self, nsteps=1):
        check(self._L.wap_step(self.handle, int(nsteps), int(self.step_index), self.t0),
              "wap_step")
        self.step_index += int(nsteps)

    def step_profiled(self, nsteps):
        """Run nsteps with per-pass HIP-event timing; returns (pass_ms[4], total_ms)."""
        pm = (ctypes.c_double * 4)()
        tot = ctypes.c_double()
        check(self._L.wap_step_profiled(self.handle, int(nsteps), int(self.step_index), self.t0,
                                        pm, ctypes.byref(tot)), "wap_step_profiled")
        self.step_index += int(nsteps)
        return [pm[i] for i in range(4)], tot.value

    def synchronize(self):
        check(self._L.wap_synchronize(self.handle), "wap_synchronize")

    def time_of_step(self, n):
        """Time at which step n injects (the reference's float or exact clock)."""
        p = self.params
        if p.time_mode == "f32":
            return float(np.float32(np.float32(n) * np.float32(p.dt)))
        return self.t0 + n * p.dt

    # --------------------------------------------------------------- output
    def get_state(self, dtype=np.complex128, device=False, stream_ordered=False):
        """psi (and pi for KG) as numpy arrays (nz,ny,nx) or, device=True, torch tensors
        (stream_ordered: filled on the solver's stream without waiting)."""
        p = self.params
        shape = p.shape
        kg = p.equation == "klein_gordon"
        if device:
            tdt = torch.complex128 if np.dtype(dtype) == np.complex128 else torch.complex64
            psi = torch.empty(shape, dtype=tdt, device="cuda")
            pi = torch.empty(shape, dtype=tdt, device="cuda") if kg else None
            code = DT_C128 if tdt == torch.complex128 else DT_C64
            check(self._L.wap_get_state(self.handle, ctypes.c_void_p(psi.data_ptr()),
                                        ctypes.c_void_p(pi.data_ptr()) if kg else None,
                                        code, WAP_STREAM_ORDERED if stream_ordered else 1),
                  "wap_get_state")
            return (psi, pi) if kg else psi
        psi = np.empty(shape, dtype=dtype)
        pi = np.empty(shape, dtype=dtype) if kg else None
        code = DT_C128 if np.dtype(dtype) == np.complex128 else DT_C64
        check(self._L.wap_get_state(self.handle, ctypes.c_void_p(psi.ctypes.data),
                                    ctypes.c_void_p(pi.ctypes.data) if kg else None, code, 0),
              "wap_get_state")
        return (psi, pi) if kg else psi

    def get_plane(self, k):
        out = np.empty((self.params.ny, self.params.nx), dtype=np.complex128)
        check(self._L.wap_get_plane(self.handle, int(k), ctypes.c_void_p(out.ctypes.data),
                                    DT_C128, 0), "wap_get_plane")
        return out

    def accumulate_plane(self, k, phase, acc):
        """acc (torch complex128 (ny,nx) on the GPU) += psi[k] e^{i phase}."""
        if not (isinstance(acc, torch.Tensor) and acc.is_cuda and acc.dtype == torch.complex128
                and acc.is_contiguous() and acc.numel() == self.params.nx * self.params.ny):
            raise ValueError("acc must be a contiguous complex128 CUDA tensor of ny*nx elements")
        check(self._L.wap_accumulate_plane(self.handle, int(k), float(np.cos(phase)),
                                           float(np.sin(phase)), ctypes.c_void_p(acc.data_ptr())),
              "wap_accumulate_plane")

    def norm2(self):
        out = ctypes.c_double()
        check(self._L.wap_norm2(self.handle, ctypes.byref(out)), "wap_norm2")
        return out.value

    def device_bytes(self):
        out = ctypes.c_int64()
        check(self._L.wap_device_bytes(self.handle, ctypes.byref(out)), "wap_device_bytes")
        return out.value

    def launch_shape(self):
        ty, tz, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(self._L.wap_launch_shape(self.handle, ctypes.byref(ty), ctypes.byref(tz),
                                       ctypes.byref(g)), "wap_launch_shape")
        return ty.value, tz.value, g.value

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._L.wap_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
