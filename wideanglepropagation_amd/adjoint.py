"""Gradient of exit-wave losses with respect to the potential V(r).

``exit_wave(V, params, nsteps, k_exit, ...)`` runs the FDTD forward on the GPU
and returns the demodulated exit wave

    E = sum_s w_s e^{+i w t_s} psi_s[k_exit]      (complex128, (ny, nx))

(the notebooks' average of demodulated exit planes,
``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:267-290``, as a linear
functional of the states psi_s after s steps, t_s = t0 + s dt).  It is a
``torch.autograd.Function``: the backward is the discrete adjoint of the
same RK4 steps (hand-written HIP kernels, ``wap_adjoint_*`` in the C ABI),
with sqrt(T) checkpointing -- states every C = ceil(sqrt(T)) steps are kept
during the forward; each segment is recomputed once with its stage inputs
recorded and then swept backwards.  Any torch loss of E (far field, FFT
intensities) composes with it.
"""
import ctypes
import math

import numpy as np
import torch

from ._lib import check, lib
from .engine import FDTD3D


class _Adjoint:
    """The C adjoint of one handle; its records live in a torch-allocated workspace, so
    that the caching allocator keeps the memory between backward passes."""

    def __init__(self, sim, slots):
        self._L = lib()
        self.sim = sim
        nb = ctypes.c_int64()
        check(self._L.wap_adjoint_workspace_bytes(sim.handle, int(slots), ctypes.byref(nb)),
              "wap_adjoint_workspace_bytes")
        self._ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        a = ctypes.c_void_p()
        check(self._L.wap_adjoint_create_in(sim.handle, int(slots),
                                            ctypes.c_void_p(self._ws.data_ptr()), int(nb.value),
                                            ctypes.byref(a)), "wap_adjoint_create_in")
        self._a = a

    def record(self, step, slot):
        check(self._L.wap_adjoint_record(self._a, int(step), self.sim.t0, int(slot)),
              "wap_adjoint_record")

    def step(self, slot):
        check(self._L.wap_adjoint_step(self._a, int(slot)), "wap_adjoint_step")

    def seed(self, k, g, c):
        check(self._L.wap_adjoint_seed(self._a, int(k), ctypes.c_void_p(g.data_ptr()),
                                       float(c.real), float(c.imag)), "wap_adjoint_seed")

    def grad(self):
        p = self.sim.params
        out = torch.empty(p.shape, dtype=torch.float64, device="cuda")
        check(self._L.wap_adjoint_get_grad(self._a, ctypes.c_void_p(out.data_ptr()), 1), "grad")
        return out

    def close(self):
        if self._a is not None:
            check(self._L.wap_adjoint_destroy(self._a), "wap_adjoint_destroy")
            self._a = None
            self._ws = None   # back to the caching allocator (destroy synchronised)


def default_samples(nsteps, n_frames=None, start_frac=0.6):
    """Sample steps and weights of the notebook recipe: frames of nsteps/n_frames steps,
    average over the last 40 % of them (frame ends)."""
    if n_frames is None:
        return {int(nsteps): 1.0}
    spf = nsteps // n_frames
    frames = list(range(int(start_frac * n_frames), n_frames))
    w = 1.0 / max(1, len(frames))
    return {(f + 1) * spf: w for f in frames}


class ExitWave(torch.autograd.Function):
    """V -> E (see module docstring).  Non-tensor arguments are passed through ``opts``."""

    @staticmethod
    def forward(ctx, V, opts):
        p = opts["params"]
        nsteps = int(opts["nsteps"])
        k_exit = int(opts["k_exit"])
        samples = opts.get("samples") or {nsteps: 1.0}
        C = int(opts.get("checkpoint_every") or max(1, math.ceil(math.sqrt(nsteps))))
        sim = FDTD3D(p, V=V.detach().contiguous(), source_profile=opts.get("source_profile"),
                     t0=opts.get("t0", 0.0))
        if opts.get("psi0") is not None:
            sim.set_state(opts["psi0"], opts.get("pi0"))
        E = torch.zeros((p.ny, p.nx), dtype=torch.complex128, device="cuda")
        tmp = torch.zeros_like(E)
        ckpt = {}
        keep = bool(ctx.needs_input_grad[0])   # no checkpoints when no gradient is wanted
        # step from event to event (a checkpoint before step n, a sample after step s)
        events = sorted({s for s in samples if 0 < s <= nsteps} |
                        ({n for n in range(0, nsteps, C)} if keep else set()) | {nsteps})
        n = 0
        for e in events:
            if keep and n % C == 0 and n not in ckpt:
                ckpt[n] = sim.get_state(np.complex128 if p.precision == "f64" else np.complex64,
                                        device=True)
            if e > n:
                sim.step(e - n)
                n = e
            if n in samples:
                tmp.zero_()
                sim.accumulate_plane(k_exit, p.e_sim * (sim.t0 + n * p.dt), tmp)
                E += samples[n] * tmp
            if keep and n % C == 0 and n < nsteps and n not in ckpt:
                ckpt[n] = sim.get_state(np.complex128 if p.precision == "f64" else np.complex64,
                                        device=True)
        if not keep:
            sim.close()
            ctx.sim = None
            ctx.ckpt = None
            return E
        ctx.sim = sim
        ctx.ckpt = ckpt
        ctx.opts = dict(opts, samples=samples, checkpoint_every=C)
        ctx.v_dtype = V.dtype
        return E

    @staticmethod
    def backward(ctx, gE):
        sim = ctx.sim
        p = sim.params
        o = ctx.opts
        nsteps, k_exit, C = int(o["nsteps"]), int(o["k_exit"]), int(o["checkpoint_every"])
        samples = o["samples"]
        g = gE.detach().to(torch.complex128).contiguous()
        adj = _Adjoint(sim, C)
        try:
            def seed(s):
                if s in samples:
                    t = sim.t0 + s * p.dt
                    c = samples[s] * complex(math.cos(p.e_sim * t), math.sin(p.e_sim * t))
                    adj.seed(k_exit, g, c)
            seed(nsteps)
            for start in reversed(range(0, nsteps, C)):
                end = min(start + C, nsteps)
                st = ctx.ckpt[start]
                if isinstance(st, tuple):
                    sim.set_state(st[0], st[1])
                else:
                    sim.set_state(st)
                for m in range(start, end):
                    adj.record(m, m - start)
                for m in reversed(range(start, end)):
                    adj.step(m - start)
                    seed(m)
            grad = adj.grad()
        finally:
            adj.close()
            sim.close()
            ctx.ckpt = None
        return grad.to(ctx.v_dtype), None


# the name the drop-in's component list gives the differentiable FDTD operator
FDTDAdjointFunction = ExitWave


def exit_wave(V, params, nsteps, k_exit, samples=None, checkpoint_every=None, psi0=None,
              pi0=None, source_profile=None, t0=0.0):
    """Differentiable exit wave E(V) (complex128 (ny, nx) on the GPU)."""
    opts = {"params": params, "nsteps": nsteps, "k_exit": k_exit, "samples": samples,
            "checkpoint_every": checkpoint_every, "psi0": psi0, "pi0": pi0,
            "source_profile": source_profile, "t0": t0}
    return ExitWave.apply(V, opts)


def far_field_intensity(E):
    """|fftshift(fft2(E))|^2 (``propagation.py:346-350``)."""
    return torch.abs(torch.fft.fftshift(torch.fft.fft2(E))) ** 2
