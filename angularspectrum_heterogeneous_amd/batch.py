"""Batched solves: many independent transmit configurations advanced together.

The reference solves one configuration per ``angular_spectrum_solve`` call
(ASS:1217); this module runs a list of (initial_field, SolverParams) instances
of uneven grid sizes side by side.  Every z-step of the batch is one
``ash_batch_step`` call: the plain steps of all active instances are captured
into one hipGraph, spread over up to four concurrent branches, and replayed
with a single graph launch.  Each instance
keeps the single-solve semantics of ``angular_spectrum_solve`` (its own dZ,
stability restart, partial last step, outputs), so its results equal a
single-instance run.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import AshStepIO, check
from .plan import MarchPlan
from .solver import SolverParams, StabilityError, _as_device_field, _Marcher, _Restart, _upload_screens


def initial_dz(field0: torch.Tensor, p: SolverParams) -> Tuple[float, float]:
    """(N, dZ) exactly as angular_spectrum_solve chooses them (ASS:1292-1296)."""
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    max_amp = float(field0.abs().max().item())
    dZ = (min(p.stabilityRecoveryFactor * p.dT / (max_amp * N + 1e-30), p.dZmin)
          if max_amp > 0 else p.dZmin)
    return N, max(dZ, p.dZmin)


class _Instance:
    def __init__(self, field, p: SolverParams, check_every: int, autotune: bool):
        self.p = p
        self.field0 = _as_device_field(field)
        self.N, dZ = initial_dz(self.field0, p)
        self.plan = MarchPlan(tuple(self.field0.shape), p, p.ash_subcycle, p.ash_cfl_target)
        if autotune:
            self.plan.autotune()
        if p.ash_learned is not None:
            self.plan.set_learned(p.ash_learned, p.ash_dpdt_max)
        self.screens = _upload_screens(p.phaseScreens)
        self.check_every = check_every
        self.start(dZ)

    def start(self, dZ: float) -> None:
        self.m = _Marcher(self.plan, self.field0, self.p, self.N, dZ, False, None, self.screens, self.check_every)


class BatchSolve:
    """Solves ``instances`` = [(initial_field, SolverParams), ...] side by side.

    ``run()`` returns, per instance, the tuple of ``angular_spectrum_solve``:
    (field, pnp, ppp, pI, pIloss, zaxis, pax).  With ``autotune`` every plan
    picks its launch configuration by timing (``MarchPlan.autotune``); the
    choice does not change the arithmetic.  ``branches`` (1-4) is the number of
    concurrent branches the instances are spread over in each step's graph.
    """

    def __init__(self, instances: Sequence[Tuple[object, SolverParams]], check_every: int = 32,
                 autotune: bool = False, release_inputs: bool = True, branches: int = 4):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.ash_batch_create(ctypes.byref(h)), "ash_batch_create")
        self.h = h
        check(self.lib.ash_batch_set_branches(h, int(branches)), "ash_batch_set_branches")
        self.items: List[_Instance] = [_Instance(f, p, check_every, autotune) for f, p in instances]
        self.release_inputs = release_inputs
        self.steps_launched = 0

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.ash_batch_destroy(self.h)
            self.h = None
        for it in getattr(self, "items", []):
            it.plan.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self) -> List[tuple]:
        results: List[Optional[tuple]] = [None] * len(self.items)
        stream = _lib.current_stream()
        while True:
            act = [i for i, it in enumerate(self.items) if results[i] is None and not it.m.done]
            for i, it in enumerate(self.items):
                if results[i] is None and it.m.done:
                    results[i] = it.m.result()
                    if self.release_inputs:
                        it.field0 = None
                        it.m = None
                        it.plan.close()
            if not act:
                break
            ios = [self.items[i].m.prepare() for i in act]
            plans = (ctypes.c_void_p * len(act))(*[self.items[i].plan.h.value for i in act])
            io_arr = (AshStepIO * len(act))(*ios)
            check(self.lib.ash_batch_step(self.h, plans, io_arr, len(act), stream), "ash_batch_step")
            self.steps_launched += 1
            for i in act:
                it = self.items[i]
                try:
                    it.m.finish()
                except _Restart as r:
                    it.start(r.dz)
        return results  # type: ignore[return-value]


def batch_solve(instances: Sequence[Tuple[object, SolverParams]], **kw) -> List[tuple]:
    """Convenience wrapper: ``BatchSolve(instances, **kw).run()``."""
    b = BatchSolve(instances, **kw)
    try:
        return b.run()
    finally:
        b.close()


__all__ = ["BatchSolve", "batch_solve", "initial_dz", "StabilityError"]
