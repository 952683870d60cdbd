"""Batched trajectories on one GPU (qlles_batch_* in include/qlles.h).

The reference advances one trajectory per process and GPU (SLURM arrays,
``slurm_*.sh:5``).  ``Batch`` advances many independent trajectories -- each
with its own plan, state and optional particles -- together, either member by
member (HIP graphs of each member's steps replayed over a few streams) or
grouped: members of one kind (dim, grid, precision, viscous mode, kernel
variants, interpolator settings) are stepped by multi-member launches, one
launch per kernel of the step for the whole group.  Each trajectory is
bit-identical to the same steps taken alone with ``_Step.advance`` /
``qlles_step_advect``.
"""

import ctypes

import torch

from . import _lib


class Batch:
    def __init__(self):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.qlles_batch_create(ctypes.byref(h)))
        self.handle = h
        self.members = []  # keeps the tensors and plans alive

    def add(self, step, u_hat, interp=None, r=None, u_r=None):
        """Add a trajectory: ``step`` from set_up_solver / rk4_wrapper, its state ``u_hat`` (updated in
        place), optionally an interpolator with positions ``r`` (np, dim) FP64 and a (np, dim) output."""
        np_ = 0
        if interp is not None:
            if r is None:
                raise ValueError("particles need positions r")
            np_ = int(r.shape[0])
            if u_r is None:
                u_r = torch.empty((np_, interp.dim), dtype=interp.dtype, device=r.device)
        idx = ctypes.c_int(-1)
        _lib.check(self._lib.qlles_batch_add(self.handle, step.plan.handle,
                                             interp.it.handle if interp is not None else None,
                                             _lib.ptr(u_hat), _lib.ptr(r) if interp is not None else None, np_,
                                             _lib.ptr(u_r) if interp is not None else None, ctypes.byref(idx)))
        self.members.append((step, u_hat, interp, r, u_r))
        return idx.value

    MODES = {False: 0, True: 1, "whole": 2, "grouped": 3, "grouped-whole": 4}

    def configure(self, lanes, chunk, single_graph=False):
        """lanes streams, chunk steps per graph; single_graph False (per-member graphs) / True (one graph per
        chunk of every member) / "whole" (one graph of every member's whole run: one launch per run) /
        "grouped" (multi-member launches per group of one kind, eager, groups interleaved chunk steps at a
        time) / "grouped-whole" (the grouped run captured into one graph: one launch per run)."""
        if single_graph not in self.MODES:
            raise ValueError(f"single_graph must be one of {list(self.MODES)}")
        _lib.check(self._lib.qlles_batch_configure(self.handle, int(lanes), int(chunk), self.MODES[single_graph]))

    def tune(self, trial_steps=16):
        """Pick streams x steps-per-graph by timing on scratch copies (trajectories do not move)."""
        _lib.check(self._lib.qlles_batch_tune(self.handle, int(trial_steps), _lib.stream_ptr()))
        return self.config()

    def config(self):
        lanes, chunk, single, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(self._lib.qlles_batch_get_config(self.handle, ctypes.byref(lanes), ctypes.byref(chunk),
                                                    ctypes.byref(single), ctypes.byref(us)))
        mode = {v: k for k, v in self.MODES.items()}[single.value]
        return dict(lanes=lanes.value, chunk=chunk.value, single_graph=mode, us_per_traj_step=us.value)

    def kinds(self):
        """The kinds the grouped modes step together: (kind of each member, time per step of each kind's
        group alone in microseconds as the tuner measured it, 0 before tuning)."""
        nk = ctypes.c_int()
        _lib.check(self._lib.qlles_batch_kinds(self.handle, ctypes.byref(nk), None, None, 0))
        kind_of = (ctypes.c_int * max(1, len(self.members)))()
        us = (ctypes.c_double * max(1, nk.value))()
        _lib.check(self._lib.qlles_batch_kinds(self.handle, ctypes.byref(nk), kind_of, us, nk.value))
        return list(kind_of)[:len(self.members)], list(us)[:nk.value]

    def run(self, nsteps):
        """Advance every member: ``nsteps`` an int (all the same) or a list (uneven lengths)."""
        if isinstance(nsteps, int):
            _lib.check(self._lib.qlles_batch_run(self.handle, None, nsteps, _lib.stream_ptr()))
        else:
            arr = (ctypes.c_int * len(nsteps))(*[int(v) for v in nsteps])
            _lib.check(self._lib.qlles_batch_run(self.handle, arr, 0, _lib.stream_ptr()))

    def __del__(self):
        try:
            if self.handle:
                self._lib.qlles_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
