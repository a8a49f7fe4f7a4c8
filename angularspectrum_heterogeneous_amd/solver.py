"""Drop-in ``SolverParams`` / ``angular_spectrum_solve`` on the gfx950 library.

Mirrors the reference driver (angular_spectrum_solver.py, ASS:1217-1567):
same signature, same step-size and crossing logic, same 7-tuple of NumPy
outputs.  The per-z march runs on the GPU as one captured hipGraph per step;
the field never returns to the host inside the loop.  Differences from the
reference, all deliberate:

* where the reference's stability restart would loop forever (the recomputed
  dZ equals the current one, ASS:1421-1439), ``StabilityError`` is raised;
  the check runs on device-recorded max|p| values every ``check_every`` steps
  (restarts are replayed exactly as the reference does when dZ does change);
* ``per_step_callback`` receives a torch tensor on the GPU instead of a JAX
  array;
* TOF extraction (``tof_ref_trace`` / ``tof_env_ratio``) runs on the GPU with
  this library's estimators: the reference imports them from a module
  (``tof_extraction``, ASS:26-29) that it does not ship, so they follow its
  docstring (ASS:1244-1264; see ``ash_plan_set_tof``); ``taxis`` must be uniform;
* ``useNonlinearityObliquity=True`` together with the learned nonlinear step
  (``ash_learned``) raises ``ValueError``;
* an initial field or source plane given as a ``SourceSlice`` (the sparse
  form of ``make_bowl_source``, ``make_bowl_source_planes`` and the sparse-array
  builder) is synthesised on the GPU instead of being uploaded dense.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from ._lib import AshStepIO
from .plan import MarchPlan
from .sources import SourceSlice


@dataclass
class SolverParams:
    """All solver parameters, same names and defaults as ASS:35-76.

    The ``ash_*`` fields are extensions of this build (the reference has no
    counterpart); their defaults reproduce the reference march exactly.
    """
    dX: float = 0.0
    dY: float = 0.0
    dT: float = 0.0
    c0: float = 1500.0
    rho0: float = 1000.0
    beta: float = 3.5
    alpha0: float = 0.5          # dB/MHz^pow/cm  (negative -> water)
    attenPow: float = 1.0
    f0: float = 5e6
    propDist: float = 0.08
    boundaryFactor: float = 0.2
    useSplitStep: bool = True
    useAdaptiveFiltering: bool = True
    useTVD: bool = True
    freqFilterThreshold: float = 0.05
    adaptiveFilterStrength: float = 0.7
    stabilityThreshold: float = 0.2
    stabilityRecoveryFactor: float = 0.15
    dZmin: float = 1e-3
    useBoundaryLayer: bool = True
    boundaryProfile: str = 'quadratic'
    useFreqWeightedBoundary: bool = False
    useSuperAbsorbing: bool = False
    superAbsorbingStrength: float = 0.8
    fluxScheme: str = 'rusanov'
    useObliquityCorrection: bool = True
    useNonlinearityObliquity: bool = False
    phaseScreens: object = None
    sourcePlanes: object = None
    # --- extensions ---
    ash_subcycle: str = 'off'       # 'off' | 'global' (_kt_flux_adaptive) | 'column': CFL sub-cycling of KT
    ash_cfl_target: float = 0.5
    ash_learned: object = None      # a learned.LearnedStep: network update on smooth columns
    ash_dpdt_max: float = 6e12      # KT on columns whose max|dp/dt| exceeds this (Pa/s)


class StabilityError(RuntimeError):
    """The stability criterion failed and dZ cannot shrink (the reference loops forever)."""


class SubcycleLimitError(RuntimeError):
    """CFL sub-cycling needed more sub-steps per dZ than the kernels take (1024);
    the reference's host loop has no bound."""


def _as_device_field(x) -> torch.Tensor:
    if isinstance(x, SourceSlice):
        return x.device_field("cuda")
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=torch.float32).contiguous()
    arr = np.ascontiguousarray(np.asarray(x), dtype=np.float32)
    return torch.from_numpy(arr).to("cuda")


def _to_host(t: torch.Tensor) -> np.ndarray:
    """Device tensor -> NumPy array through page-locked host memory (a pageable
    copy runs at a tenth of the link rate).  The array is a view of the pinned
    buffer, which lives as long as the array does."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy()


class _Restart(Exception):
    def __init__(self, dz):
        self.dz = dz


def angular_spectrum_solve(
        initial_field,
        params: SolverParams,
        verbose: bool = True,
        per_step_callback: Optional[Callable[[int, float, torch.Tensor], None]] = None,
        tof_ref_trace: Optional[np.ndarray] = None,
        tof_t_ref_peak_s: Optional[float] = None,
        tof_env_ratio: Optional[float] = None,
        taxis: Optional[np.ndarray] = None,
        check_every: int = 32,
        _plan_out: Optional[list] = None):
    """Propagates a 3-D field with the modified angular spectrum method.

    Same contract as ASS:1217-1567; returns
    ``(field, pnp, ppp, pI, pIloss, zaxis, pax)`` as NumPy arrays:
    field (nX,nY,nT) float32; pnp/ppp/pI/pIloss (nX,nY,nZ) float32;
    zaxis (nZ,) float64; pax (nT,nZ) float32; with TOF enabled (``tof_ref_trace``,
    which takes precedence, or ``tof_env_ratio``) an eighth output, tof
    (nX,nY,nZ) float32, the arrival time (retarded time + z / c0) per column.
    """
    tof_cfg = None
    if tof_ref_trace is not None or tof_env_ratio is not None:
        if tof_ref_trace is not None and tof_t_ref_peak_s is None:
            raise ValueError('tof_t_ref_peak_s is required when tof_ref_trace is provided')
        tof_cfg = (tof_ref_trace, tof_t_ref_peak_s, tof_env_ratio, taxis)
    if params.useNonlinearityObliquity and params.ash_learned is not None:
        raise ValueError('the learned nonlinear step does not take the beam obliquity scale')
    if params.ash_subcycle != 'off' and params.ash_learned is not None:
        raise ValueError('the learned nonlinear step replaces the un-sub-cycled KT update only')

    field0 = _as_device_field(initial_field)
    nX, nY, nT = field0.shape
    c0 = params.c0
    N = params.beta / (2 * c0 ** 3 * params.rho0)
    max_amp = float(field0.abs().max().item())
    dZ = (min(params.stabilityRecoveryFactor * params.dT / (max_amp * N + 1e-30), params.dZmin)
          if max_amp > 0 else params.dZmin)
    dZ = max(dZ, params.dZmin)

    plan = MarchPlan((nX, nY, nT), params, params.ash_subcycle, params.ash_cfl_target)
    if tof_cfg is not None:
        _configure_tof(plan, tof_cfg, nT, params.dT)
    if params.ash_learned is not None:
        plan.set_learned(params.ash_learned, params.ash_dpdt_max)
    if _plan_out is not None:
        _plan_out.append(plan)
    screens = _upload_screens(params.phaseScreens)
    while True:
        try:
            return _march(plan, field0, params, N, dZ, verbose, per_step_callback, screens, check_every,
                          tof_cfg is not None)
        except _Restart as r:
            if verbose:
                print('Stability criterion violated — reducing step size')
            dZ = r.dz


def _configure_tof(plan: MarchPlan, cfg, nT: int, dT: float) -> None:
    """TOF on the plan (ASS:1387-1397): matched filter when a reference trace is
    given, else the envelope threshold; the time axis must be uniform."""
    ref, t_ref_peak, env_ratio, taxis = cfg
    tax = np.asarray(taxis, dtype=np.float64) if taxis is not None else np.arange(nT, dtype=np.float64) * dT
    if tax.shape != (nT,):
        raise ValueError(f'taxis must have nT = {nT} samples')
    dt = float(tax[1] - tax[0]) if nT > 1 else float(dT)
    if nT > 2 and np.max(np.abs(np.diff(tax) - dt)) > 1e-9 * abs(dt):
        raise ValueError('TOF extraction needs a uniform time axis')
    if ref is not None:
        r = np.asarray(ref, dtype=np.float32)
        if r.shape != (nT,):
            raise ValueError(f'tof_ref_trace must have nT = {nT} samples')
        plan.set_tof(_lib.TOF_MATCHED, 1.0, float(tax[0]), dt, torch.from_numpy(r).to('cuda'),
                     float(t_ref_peak))
    else:
        plan.set_tof(_lib.TOF_ENVELOPE, float(env_ratio), float(tax[0]), dt)


def _upload_screens(screens):
    if screens is None:
        return None
    out = []
    for scr in screens:
        ph = torch.from_numpy(np.ascontiguousarray(np.asarray(scr[1]), dtype=np.float32)).to("cuda")
        amp = None
        if len(scr) > 2 and scr[2] is not None:
            amp = torch.from_numpy(np.ascontiguousarray(np.asarray(scr[2]), dtype=np.float32)).to("cuda")
        out.append((scr[0], ph, amp))
    return out


class _Marcher:
    """One solve's z loop as a state machine (ASS:1413-1550): ``prepare`` fills the
    step io (screens and sources crossed in the step), the caller launches the
    step, ``finish`` books it (deferred stability check every ``check_every``
    steps, callback) and ``result`` returns the outputs.  Used by the single
    solve and by the batched solve."""

    def __init__(self, plan: MarchPlan, field0, p: SolverParams, N, dZ, verbose, cb, screens, check_every,
                 tof: bool = False):
        nX, nY, nT = field0.shape
        self.plan, self.p, self.N, self.dZ = plan, p, N, dZ
        self.verbose, self.cb, self.screens, self.check_every = verbose, cb, screens, check_every
        self.dev = dev = field0.device
        plan.set_dz(dZ)
        self.max_steps = max_steps = int(np.ceil(p.propDist / dZ)) + 100
        self.outs = torch.zeros((4, max_steps, nX, nY), dtype=torch.float32, device=dev)
        self.pax = torch.zeros((max_steps, nT), dtype=torch.float32, device=dev)
        self.stats = torch.zeros(max_steps + 1, dtype=torch.int32, device=dev)
        self.field_cb = torch.empty((nX, nY, nT), dtype=torch.float32, device=dev) if cb is not None else None
        self.tof = torch.zeros((max_steps, nX, nY), dtype=torch.float32, device=dev) if tof else None
        plan.load(field0, self.stats)
        io = self.io = AshStepIO()
        io.max_rows = max_steps
        io.pnp_dev, io.ppp_dev, io.pI_dev, io.pIloss_dev = (self.outs[k].data_ptr() for k in range(4))
        io.pax_dev = self.pax.data_ptr()
        io.stats_dev = self.stats.data_ptr()
        io.field_out_dev = self.field_cb.data_ptr() if self.field_cb is not None else None
        io.tof_dev = self.tof.data_ptr() if self.tof is not None else None
        self.keep = []      # device copies of sources crossed in the current step
        self.zvec = []
        self.z = 0.0        # running sum(zvec) (ASS:1413-1415 recomputes it every step)
        self.dz_steps = []  # dZ_step of each step (stability check)
        self.checked = 0
        self.cc = 0

    @property
    def done(self) -> bool:
        return self.z >= self.p.propDist - 1e-15

    def _check_stability(self, upto):
        if upto <= self.checked:
            return
        if self.plan.desc.subcycle != _lib.SUBCYCLE_OFF:
            n_capped = self.plan.subcycle_capped()
            if n_capped:
                raise SubcycleLimitError(
                    f'CFL sub-cycling needed more than {_lib.MAX_SUBSTEPS} sub-steps per dZ in '
                    f'{n_capped} column(s) or field(s) by step {upto - 1}')
        p, N = self.p, self.N
        mx = self.stats[self.checked:upto].cpu().view(torch.float32).numpy().astype(np.float64)
        for k in range(upto - self.checked):
            step = self.checked + k
            if N * self.dz_steps[step] / p.dT * mx[k] > p.stabilityThreshold:
                new_dz = max(p.stabilityRecoveryFactor * p.dT / (mx[k] * N), p.dZmin)
                if new_dz == self.dZ:
                    raise StabilityError(
                        f'stability criterion violated at step {step} (max|p|={mx[k]:.4g}) and dZ '
                        f'cannot shrink below dZmin={p.dZmin}')
                raise _Restart(new_dz)
        self.checked = upto

    def prepare(self) -> AshStepIO:
        p, io, cc = self.p, self.io, self.cc
        if cc >= self.max_steps:
            raise RuntimeError('step count exceeded the output allocation')
        z_cur = self.z
        remaining = p.propDist - z_cur
        dZ_step = min(self.dZ, remaining)
        self.dz_steps.append(dZ_step)
        if self.verbose:
            print(f'z = {z_cur + dZ_step:.6f} m  (step {cc})')
        io.row = cc
        io.dz = dZ_step
        # screens crossed in this step (ASS:1505-1518)
        io.n_screens = 0
        if self.screens is not None:
            for z_ps, ph, amp in self.screens:
                if z_cur < z_ps <= z_cur + dZ_step:
                    if io.n_screens == 4:
                        raise NotImplementedError('more than 4 screens in one step')
                    io.screen_phase_dev[io.n_screens] = ph.data_ptr()
                    io.screen_amp_dev[io.n_screens] = amp.data_ptr() if amp is not None else None
                    io.n_screens += 1
                    if self.verbose:
                        print(f'  applied phase screen at z = {z_ps:.4f} m')
        # sources crossed in this step (ASS:1521-1527)
        keep = self.keep
        keep.clear()
        io.n_src_dense = 0
        io.src_mask_dev = None
        if p.sourcePlanes is not None:
            sparse = []
            for sp_z, sp in p.sourcePlanes:
                if z_cur < sp_z <= z_cur + dZ_step:
                    if isinstance(sp, SourceSlice):
                        sparse.append(sp)
                    else:
                        if io.n_src_dense == 4:
                            raise NotImplementedError('more than 4 dense source planes in one step')
                        t = _as_device_field(sp)
                        keep.append(t)
                        io.src_dense_dev[io.n_src_dense] = t.data_ptr()
                        io.n_src_dense += 1
                    if self.verbose:
                        print(f'  injected source plane at z = {sp_z*1e3:.2f} mm')
            if sparse:
                merged = SourceSlice.merge(sparse)
                mask, delay, tax = merged.device_arrays(self.dev)
                keep.extend([mask, delay, tax])
                io.src_mask_dev = mask.data_ptr()
                io.src_delay_dev = delay.data_ptr()
                io.taxis_dev = tax.data_ptr()
                io.src_f0, io.src_p0 = merged.f0, merged.p0
                io.src_ncycles, io.src_dur = merged.ncycles, float(merged.dur)
        return io

    def finish(self) -> None:
        dZ_step = self.dz_steps[-1]
        self.zvec.append(dZ_step)
        self.z = sum(self.zvec)
        if self.keep:
            torch.cuda.current_stream().synchronize()
        if self.cb is not None:
            self._check_stability(self.cc + 1)
            self.cb(self.cc, self.z, self.field_cb)
        self.cc += 1
        if self.cc % self.check_every == 0:
            self._check_stability(self.cc)

    def result(self):
        cc = self.cc
        self._check_stability(cc)
        field = self.plan.store()
        res = _to_host(self.outs[:, :cc].permute(0, 2, 3, 1).contiguous())
        pax_np = _to_host(self.pax[:cc].t().contiguous())
        zaxis = np.cumsum(self.zvec[:cc])
        out = (_to_host(field), res[0], res[1], res[2], res[3], zaxis, pax_np)
        if self.tof is not None:
            # retarded arrival time + z / c0 (the march's time axis travels with the pulse)
            tof = _to_host(self.tof[:cc].permute(1, 2, 0).contiguous()).astype(np.float64)
            out += ((tof + zaxis[None, None, :] / self.p.c0).astype(np.float32),)
        return out


def _march(plan: MarchPlan, field0, p: SolverParams, N, dZ, verbose, cb, screens, check_every, tof=False):
    m = _Marcher(plan, field0, p, N, dZ, verbose, cb, screens, check_every, tof)
    while not m.done:
        io = m.prepare()
        plan.step(io)
        m.finish()
    return m.result()
