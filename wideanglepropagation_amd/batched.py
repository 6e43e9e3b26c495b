"""Many independent grids stepped by one launch per pass, optionally as one hipGraph.

The reference runs one grid per process (``fdtd_3d.cu:275-403``); a set of
probe positions, beam energies or wave packets is a set of independent runs.
``BatchedFDTD`` groups ``FDTD3D`` handles by (equation, precision, schedule)
and, per group, issues one kernel launch per pass whose workgroups are the
union of every grid's column tiles (equal-sized work units, so uneven grids
balance by construction).  ``capture(n)`` records n steps of every group into
one hipGraph (several groups, e.g. the two equations of a mixed workload, run on
forked streams inside it, so their launches overlap); ``replay()`` rewrites the
per-step source phasors and launches it.
"""
import ctypes

import torch

from ._lib import check, lib


class BatchedFDTD:
    """One batched launch group per (equation, precision, schedule) of the given grids."""

    def __init__(self, sims, tile_y=0):
        if not sims:
            raise ValueError("need at least one FDTD3D")
        self._L = lib()
        self.sims = list(sims)
        groups = {}
        for s in self.sims:
            p = s.params
            groups.setdefault((p.equation, p.precision, p.resolved_schedule()), []).append(s)
        self.groups = []
        self.stream = torch.cuda.current_stream()
        for key, members in groups.items():
            arr = (ctypes.c_void_p * len(members))(*[m.handle.value for m in members])
            b = ctypes.c_void_p()
            # fused groups keep their handles' workgroup shape (64 x 4, 8 or 16 waves; the
            # one-pass Schrodinger shape is fixed)
            ty = 0 if key[2] in ("kg_fused", "schr_fused", "schr_1pass") else int(tile_y)
            check(self._L.wap_batch_create(arr, len(members), ty, ctypes.byref(b)),
                  "wap_batch_create")
            self.groups.append((key, members, b))
        self.step_index = 0
        self.captured_steps = 0
        self._set = None   # the graph over all groups (wap_batch_set) when there are several

    @property
    def voxels(self):
        return sum(s.params.voxels for s in self.sims)

    def tiles(self):
        out = []
        for key, _, b in self.groups:
            t, ty = ctypes.c_int64(), ctypes.c_int32()
            check(self._L.wap_batch_tiles(b, ctypes.byref(t), ctypes.byref(ty)), "tiles")
            out.append((key, t.value, ty.value))
        return out

    def step(self, nsteps=1):
        for _, members, b in self.groups:
            check(self._L.wap_batch_step(b, int(nsteps), int(self.step_index)), "wap_batch_step")
        self.step_index += int(nsteps)
        for s in self.sims:
            s.step_index = self.step_index

    def capture(self, nsteps):
        """Record nsteps steps of every group into one hipGraph."""
        if len(self.groups) == 1:
            check(self._L.wap_batch_capture(self.groups[0][2], int(nsteps)), "wap_batch_capture")
        else:
            if self._set is None:
                arr = (ctypes.c_void_p * len(self.groups))(*[b.value for _, _, b in self.groups])
                h = ctypes.c_void_p()
                check(self._L.wap_batch_set_create(arr, len(self.groups), ctypes.byref(h)),
                      "wap_batch_set_create")
                self._set = h
            check(self._L.wap_batch_set_capture(self._set, int(nsteps)), "wap_batch_set_capture")
        self.captured_steps = int(nsteps)

    def replay(self, timed=False):
        """Run the captured steps from the current step index; returns device ms if timed."""
        if not self.captured_steps:
            raise RuntimeError("capture() first")
        total = 0.0
        targets = ([(self._L.wap_batch_set_replay, self._set)] if self._set is not None else
                   [(self._L.wap_batch_replay, b) for _, _, b in self.groups])
        for fn, h in targets:
            ms = ctypes.c_double()
            check(fn(h, int(self.step_index), ctypes.byref(ms) if timed else None), "replay")
            total += ms.value
        self.step_index += self.captured_steps
        for s in self.sims:
            s.step_index = self.step_index
        return total if timed else None

    def synchronize(self):
        if self._set is not None:
            check(self._L.wap_batch_set_synchronize(self._set), "wap_batch_set_synchronize")
        for _, _, b in self.groups:
            check(self._L.wap_batch_synchronize(b), "wap_batch_synchronize")

    def close(self):
        if self._set is not None:
            self._L.wap_batch_set_destroy(self._set)
            self._set = None
        for _, _, b in self.groups:
            if b:
                self._L.wap_batch_destroy(b)
        self.groups = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
