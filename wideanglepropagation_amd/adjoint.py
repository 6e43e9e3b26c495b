"""Gradient of exit-wave losses with respect to the potential V(r).

``exit_wave(V, params, nsteps, k_exit, ...)`` runs the FDTD forward on the GPU
and returns the demodulated exit wave

    E = sum_s w_s e^{+i w t_s} psi_s[k_exit]      (complex128, (ny, nx))

(the notebooks' average of demodulated exit planes,
``FDTDmethod/fdtd_3D_baseline_comparison.ipynb:267-290``, as a linear
functional of the states psi_s after s steps, t_s the solver's clock:
t0 + s dt, or the reference's float32 time).  With ``n_frames`` the weights
w_s are the recipe's window (equal weights on the last 40 % of the frame ends);
with ``k_monitor`` E is divided by sqrt(mean |M|^2), M the same average of plane
k_monitor, when that exceeds 1e-9 (the recipe's incident-intensity
normalisation; differentiable, since M is another plane of the same forward).
``ExitWave`` is a ``torch.autograd.Function`` of V returning such averages for a
list of planes: the backward is the discrete adjoint of the
same RK4 steps (hand-written HIP kernels, ``wap_adjoint_*`` in the C ABI),
with checkpointing: the T steps are cut into segments of S >= ceil(sqrt(T))
steps, counted back from step T.  The forward keeps the state at the start of
every segment but the last and records the last segment's steps as it runs
them (a record step is the forward step with its output kept); the backward
sweeps the last segment straight away and recomputes each earlier one once,
with its records, before sweeping it.  S is the longest segment whose records
and checkpoints fit a memory budget (half the free device memory unless
``memory_budget`` says otherwise), so the backward recomputes T - S steps:
none when the whole run fits, as it does for grids of a few million voxels
over thousands of steps in 288 GB.  ``checkpoint_every=C`` fixes S = C.  Any
torch loss of E (far field, FFT intensities) composes with it.
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
        self.sim = sim   # held, so that the handle outlives this object
        self._a = None
        nb = self.workspace_bytes(sim, slots)
        self._ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        a = ctypes.c_void_p()
        check(self._L.wap_adjoint_create_in(sim.handle, int(slots),
                                            ctypes.c_void_p(self._ws.data_ptr()), nb,
                                            ctypes.byref(a)), "wap_adjoint_create_in")
        self._a = a

    @staticmethod
    def workspace_bytes(sim, slots):
        nb = ctypes.c_int64()
        check(lib().wap_adjoint_workspace_bytes(sim.handle, int(slots), ctypes.byref(nb)),
              "wap_adjoint_workspace_bytes")
        return int(nb.value)

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
        # before the handle is destroyed: the records swap buffers with the handle, and
        # destroy gives the handle its own buffers back
        if self._a is not None:
            check(self._L.wap_adjoint_destroy(self._a), "wap_adjoint_destroy")
            self._a = None
            self._ws = None   # back to the caching allocator (destroy synchronised)

    def __del__(self):   # a forward whose backward never ran
        try:
            self.close()
        except Exception:
            pass


def default_samples(nsteps, n_frames=None, start_frac=0.6):
    """Sample steps and weights of the notebook recipe: frames of nsteps/n_frames steps,
    average over the last 40 % of them (frame ends)."""
    if n_frames is None:
        return {int(nsteps): 1.0}
    spf = nsteps // n_frames
    frames = list(range(int(start_frac * n_frames), n_frames))
    w = 1.0 / max(1, len(frames))
    return {(f + 1) * spf: w for f in frames}


# the plan of the most recent differentiable forward (for reports): segment length,
# checkpoints kept, steps the backward recomputes
last_plan = {}


def plan_segment(nsteps, slot_bytes, ckpt_bytes, fixed_bytes, budget, min_len):
    """Segment length S: the longest in [min_len, T] whose S record slots, ceil(T/S) - 1
    checkpoints and the adjoint's fixed buffers fit in budget bytes (min_len when none
    does)."""
    T = int(nsteps)
    lo = max(1, min(int(min_len), T))
    hi = T if slot_bytes <= 0 else max(lo, min(T, int((budget - fixed_bytes) // slot_bytes)))
    for S in range(hi, lo - 1, -1):
        if S * slot_bytes + (-(-T // S) - 1) * ckpt_bytes + fixed_bytes <= budget:
            return S
    return lo


def segment_starts(nsteps, S):
    """First steps of the segments of S steps counted back from nsteps (ascending; the first
    segment is the short one)."""
    return sorted({max(0, int(nsteps) - j * S) for j in range(1, -(-int(nsteps) // S) + 1)})


def _default_budget():
    free, _ = torch.cuda.mem_get_info()
    cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
    return (free + cached) // 2


def _segment_length(sim, nsteps, opts):
    C = opts.get("checkpoint_every")
    if C:
        return max(1, min(int(C), int(nsteps)))
    p = sim.params
    w1 = _Adjoint.workspace_bytes(sim, 1)
    slot = _Adjoint.workspace_bytes(sim, 2) - w1
    ncomp = 2 if p.equation == "klein_gordon" else 1
    ckpt = p.voxels * ncomp * (16 if p.precision == "f64" else 8)
    budget = opts.get("memory_budget")
    budget = _default_budget() if budget is None else int(budget)
    return plan_segment(nsteps, slot, ckpt, w1 - slot, budget,
                        max(1, math.ceil(math.sqrt(nsteps))))


class ExitWave(torch.autograd.Function):
    """V -> E (see module docstring).  Non-tensor arguments are passed through ``opts``."""

    @staticmethod
    def forward(ctx, V, opts):
        p = opts["params"]
        nsteps = int(opts["nsteps"])
        # k_exit: one plane (E is (ny, nx)) or a list of planes (E is (len, ny, nx))
        planes = opts["k_exit"]
        single = not isinstance(planes, (list, tuple))
        planes = [int(planes)] if single else [int(k) for k in planes]
        if nsteps < 1:
            raise ValueError(f"exit_wave: nsteps must be >= 1 (got {nsteps})")
        if not planes or any(k < 0 or k >= p.nz for k in planes):
            raise ValueError(f"exit_wave: planes {planes} outside 0..{p.nz - 1}")
        samples = opts.get("samples") or {nsteps: 1.0}
        sim = FDTD3D(p, V=V.detach().contiguous(), source_profile=opts.get("source_profile"),
                     t0=opts.get("t0", 0.0))
        if opts.get("psi0") is not None:
            sim.set_state(opts["psi0"], opts.get("pi0"))
        E = torch.zeros((len(planes), p.ny, p.nx), dtype=torch.complex128, device="cuda")
        tmp = torch.zeros_like(E[0])
        ckpt = {}
        keep = bool(ctx.needs_input_grad[0])   # no checkpoints when no gradient is wanted
        cdt = np.complex128 if p.precision == "f64" else np.complex64

        def sample(n):
            if n in samples:
                for i, k in enumerate(planes):
                    tmp.zero_()
                    sim.accumulate_plane(k, p.e_sim * sim.time_of_step(n), tmp)
                    E[i].add_(samples[n] * tmp)

        adj = None
        try:
            if keep:
                S = _segment_length(sim, nsteps, opts)
                starts = segment_starts(nsteps, S)
                last = starts[-1]
                last_plan.clear()
                last_plan.update(segment=S, checkpoints=len(starts) - 1, recomputed_steps=last)
            else:
                starts, last = [], nsteps
            # plain steps up to the last segment, from event to event (a checkpoint at a
            # segment start, a sample after step s)
            # (a sample at step 0 is the initial state, with or without a gradient)
            events = sorted({s for s in samples if 0 <= s <= last} | set(starts[:-1]) | {last})
            n = 0
            for e in events:
                if e > n:
                    sim.step(e - n)
                    n = e
                sample(n)
                if n in starts[:-1]:   # stream-ordered: no host wait per checkpoint
                    ckpt[n] = sim.get_state(cdt, device=True, stream_ordered=True)
            if keep:   # the last segment: each step recorded for the backward
                adj = _Adjoint(sim, nsteps - last)
                for m in range(last, nsteps):
                    adj.record(m, m - last)
                    sample(m + 1)
        except BaseException:
            if adj is not None:
                adj.close()
            sim.close()
            raise
        ctx.single = single
        if not keep:
            sim.close()
            ctx.sim = None
            ctx.ckpt = None
            return E[0] if single else E
        ctx.sim = sim
        ctx.adj = adj
        ctx.ckpt = ckpt
        ctx.starts = starts
        ctx.opts = dict(opts, samples=samples, planes=planes)
        ctx.v_dtype = V.dtype
        return E[0] if single else E

    @staticmethod
    def backward(ctx, gE):
        sim = ctx.sim
        if sim is None:
            raise RuntimeError("exit_wave: backward called twice")
        p = sim.params
        o = ctx.opts
        nsteps, planes = int(o["nsteps"]), o["planes"]
        samples = o["samples"]
        starts = ctx.starts
        g = gE.detach().to(torch.complex128).reshape(len(planes), p.ny, p.nx).contiguous()
        adj = ctx.adj
        try:
            def seed(s):
                if s in samples:
                    t = sim.time_of_step(s)
                    c = samples[s] * complex(math.cos(p.e_sim * t), math.sin(p.e_sim * t))
                    for i, k in enumerate(planes):
                        adj.seed(k, g[i], c)
            seed(nsteps)
            ends = starts[1:] + [nsteps]
            for i in reversed(range(len(starts))):
                start, end = starts[i], ends[i]
                if i < len(starts) - 1:   # recompute (the last segment was recorded forward)
                    # the copy runs on the solver's stream (torch's stream when the
                    # solver was made), so the checkpoint's memory may go back to torch's
                    # allocator right away
                    st = ctx.ckpt.pop(start)
                    if isinstance(st, tuple):
                        sim.set_state(st[0], st[1], stream_ordered=True)
                    else:
                        sim.set_state(st, stream_ordered=True)
                    for m in range(start, end):
                        adj.record(m, m - start)
                for m in reversed(range(start, end)):
                    adj.step(m - start)
                    seed(m)
            grad = adj.grad()
        finally:
            adj.close()
            sim.close()
            ctx.adj = None
            ctx.sim = None
            ctx.ckpt = None
        return grad.to(ctx.v_dtype), None


# the name the drop-in's component list gives the differentiable FDTD operator
FDTDAdjointFunction = ExitWave


def exit_wave(V, params, nsteps, k_exit, samples=None, checkpoint_every=None, psi0=None,
              pi0=None, source_profile=None, t0=0.0, memory_budget=None, k_monitor=None,
              n_frames=None, start_frac=0.6):
    """Differentiable exit wave E(V) (complex128 (ny, nx) on the GPU).

    samples: {step: weight} of the average; without it, n_frames gives the recipe's
    window (default_samples(nsteps, n_frames, start_frac)), and without either E is the
    demodulated plane after the last step.  k_monitor: divide by the incident amplitude
    sqrt(mean |M|^2) of that plane's average when its intensity exceeds 1e-9.
    memory_budget caps the device bytes of the backward's records and checkpoints."""
    if samples is None and n_frames is not None:
        samples = default_samples(int(nsteps), int(n_frames), start_frac)
    planes = int(k_exit) if k_monitor is None else [int(k_exit), int(k_monitor)]
    opts = {"params": params, "nsteps": nsteps, "k_exit": planes, "samples": samples,
            "checkpoint_every": checkpoint_every, "psi0": psi0, "pi0": pi0,
            "source_profile": source_profile, "t0": t0, "memory_budget": memory_budget}
    E = ExitWave.apply(V, opts)
    if k_monitor is None:
        return E
    inc = torch.mean(torch.abs(E[1]) ** 2)
    if float(inc.detach()) > 1e-9:
        return E[0] / torch.sqrt(inc)
    return E[0]


def far_field_intensity(E):
    """|fftshift(fft2(E))|^2 (``propagation.py:346-350``)."""
    return torch.abs(torch.fft.fftshift(torch.fft.fft2(E))) ** 2
