"""Python handle of a gfx950 march plan (C ABI ``ash_plan``) on PyTorch tensors."""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from ._lib import AshDesc, AshStepIO, check, ptr

_PROFILES = {"quadratic": _lib.PROFILE_QUADRATIC, "wendland": _lib.PROFILE_WENDLAND}


def desc_from_params(shape, params, subcycle: str = "off", cfl_target: float = 0.5) -> AshDesc:
    """Fills the C descriptor from a SolverParams-like object (ASS:35-76)."""
    nX, nY, nT = (int(s) for s in shape)
    if params.boundaryProfile not in _PROFILES:
        # the reference treats every value other than 'wendland' as quadratic (ASS:117)
        prof = _lib.PROFILE_QUADRATIC
    else:
        prof = _PROFILES[params.boundaryProfile]
    d = AshDesc()
    d.nX, d.nY, d.nT = nX, nY, nT
    d.dX, d.dY, d.dT = float(params.dX), float(params.dY), float(params.dT)
    d.c0, d.rho0, d.beta = float(params.c0), float(params.rho0), float(params.beta)
    d.alpha0, d.attenPow, d.f0 = float(params.alpha0), float(params.attenPow), float(params.f0)
    d.boundaryFactor = float(params.boundaryFactor)
    d.useSplitStep = int(bool(params.useSplitStep))
    d.useAdaptiveFiltering = int(bool(params.useAdaptiveFiltering))
    d.useTVD = int(bool(params.useTVD))
    d.freqFilterThreshold = float(params.freqFilterThreshold)
    d.adaptiveFilterStrength = float(params.adaptiveFilterStrength)
    d.useBoundaryLayer = int(bool(params.useBoundaryLayer))
    d.boundaryProfile = prof
    d.useFreqWeightedBoundary = int(bool(params.useFreqWeightedBoundary))
    d.useSuperAbsorbing = int(bool(params.useSuperAbsorbing))
    d.superAbsorbingStrength = float(params.superAbsorbingStrength)
    d.fluxScheme = 1 if params.fluxScheme == "kt" else 0
    d.useObliquityCorrection = int(bool(params.useObliquityCorrection))
    modes = {"off": _lib.SUBCYCLE_OFF, "column": _lib.SUBCYCLE_COLUMN, "global": _lib.SUBCYCLE_GLOBAL}
    if subcycle not in modes:
        raise ValueError(f"ash_subcycle must be one of {sorted(modes)}, got {subcycle!r}")
    if subcycle != "off" and not (params.useSplitStep and params.fluxScheme == "kt"):
        raise ValueError("CFL sub-cycling applies to the split-step KT update only (fluxScheme='kt')")
    d.subcycle = modes[subcycle]
    d.cflTarget = float(cfl_target)
    d.useNonlinearityObliquity = int(bool(params.useNonlinearityObliquity))
    return d


class MarchPlan:
    """Owns one ``ash_plan``: the spectral field state and the step graph."""

    def __init__(self, shape, params, subcycle: str = "off", cfl_target: float = 0.5):
        self.lib = _lib.load()
        self.shape = tuple(int(s) for s in shape)
        self.desc = desc_from_params(self.shape, params, subcycle, cfl_target)
        h = ctypes.c_void_p()
        check(self.lib.ash_plan_create(ctypes.byref(self.desc), ctypes.byref(h)), "ash_plan_create")
        self.h = h
        self.dz: Optional[float] = None

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.ash_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- march ---------------------------------------------------------------
    def set_learned(self, net, dpdt_max: float) -> None:
        """Network update (a learned.LearnedStep, or None for KT) on columns with max|dp/dt| <= dpdt_max (Pa/s)."""
        self.net = net
        check(self.lib.ash_plan_set_learned(self.h, net.h if net is not None else None, float(dpdt_max)),
              "ash_plan_set_learned")

    def set_tof(self, mode: int, env_ratio: float, t0: float, dt: float, ref=None, t_ref_peak: float = 0.0):
        """Time of flight after every step (ash_plan_set_tof); ref: (nT,) float32 CUDA tensor."""
        self._tof_ref = ref
        check(self.lib.ash_plan_set_tof(self.h, int(mode), float(env_ratio), float(t0), float(dt), ptr(ref),
                                        float(t_ref_peak), _lib.current_stream()), "ash_plan_set_tof")

    def autotune(self) -> dict:
        """Times the radix orders and the column grid on this plan and keeps the fastest."""
        ch = (ctypes.c_int32 * 3)()
        check(self.lib.ash_plan_autotune(self.h, ch, _lib.current_stream()), "ash_plan_autotune")
        self.tuned = {"x_radix_order": int(ch[0]), "y_radix_order": int(ch[1]), "column_grid": int(ch[2])}
        return self.tuned

    def set_launch(self, x_radix_order: int = 0, y_radix_order: int = 0, column_grid: int = 2048) -> None:
        """Pins the launch configuration (the autotuner's choices)."""
        ch = (ctypes.c_int32 * 3)(x_radix_order, y_radix_order, column_grid)
        check(self.lib.ash_plan_set_launch(self.h, ch), "ash_plan_set_launch")

    def set_dz(self, dz: float) -> None:
        check(self.lib.ash_plan_set_dz(self.h, float(dz)), "ash_plan_set_dz")
        self.dz = float(dz)

    def load(self, field: torch.Tensor, stats: Optional[torch.Tensor] = None) -> None:
        _check_field(field, self.shape)
        check(self.lib.ash_load_field(self.h, ptr(field), ptr(stats), _lib.current_stream()),
              "ash_load_field")

    def store(self, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float32, device="cuda")
        _check_field(out, self.shape)
        check(self.lib.ash_store_field(self.h, ptr(out), _lib.current_stream()), "ash_store_field")
        return out

    def step(self, io: AshStepIO) -> None:
        check(self.lib.ash_step(self.h, ctypes.byref(io), _lib.current_stream()), "ash_step")

    def workspace_bytes(self) -> int:
        return int(self.lib.ash_plan_workspace_bytes(self.h))

    # ---- operator-level entry points (real field in, real field out) ---------
    def _io(self, x: torch.Tensor, out: Optional[torch.Tensor]):
        _check_field(x, self.shape)
        if out is None:
            out = torch.empty_like(x)
        _check_field(out, self.shape)
        return out

    def march_step(self, x, out=None):
        out = self._io(x, out)
        check(self.lib.ash_march_step(self.h, ptr(x), ptr(out), _lib.current_stream()), "ash_march_step")
        return out

    def attenuation_step(self, x, scale: float, out=None):
        out = self._io(x, out)
        check(self.lib.ash_attenuation_step(self.h, ptr(x), ptr(out), float(scale), _lib.current_stream()),
              "ash_attenuation_step")
        return out

    def angular_spectrum_step(self, x, half: bool, out=None):
        out = self._io(x, out)
        check(self.lib.ash_angular_spectrum_step(self.h, ptr(x), ptr(out), int(bool(half)),
                                                 _lib.current_stream()), "ash_angular_spectrum_step")
        return out

    def nonlinear_step(self, x, scheme: int, dz: float, beta_tvd: float = 1.0, subcycle: int = 0, out=None):
        if subcycle != _lib.SUBCYCLE_OFF and scheme not in (_lib.NL_KT, _lib.NL_KT_MINMOD):
            raise ValueError("CFL sub-cycling applies to the KT schemes only")
        out = self._io(x, out)
        check(self.lib.ash_nonlinear_step(self.h, ptr(x), ptr(out), int(scheme), float(dz), float(beta_tvd),
                                          int(subcycle), _lib.current_stream()), "ash_nonlinear_step")
        return out

    def subcycle_capped(self) -> int:
        """Columns (or fields) whose sub-cycling hit the kernels' bound of
        MAX_SUBSTEPS sub-steps since the last load (synchronises)."""
        n = ctypes.c_int64()
        check(self.lib.ash_plan_subcycle_capped(self.h, ctypes.byref(n)), "ash_plan_subcycle_capped")
        return int(n.value)

    def max_abs(self, x) -> float:
        _check_field(x, self.shape)
        bits = torch.zeros(1, dtype=torch.int32, device=x.device)
        check(self.lib.ash_max_abs(self.h, ptr(x), ptr(bits), _lib.current_stream()), "ash_max_abs")
        return float(bits.view(torch.float32).item())

    def freq_weighted_boundary_step(self, x, out=None):
        out = self._io(x, out)
        check(self.lib.ash_freq_weighted_boundary_step(self.h, ptr(x), ptr(out), _lib.current_stream()),
              "ash_freq_weighted_boundary_step")
        return out

    def super_absorbing_step(self, x, out=None):
        out = self._io(x, out)
        check(self.lib.ash_super_absorbing_step(self.h, ptr(x), ptr(out), _lib.current_stream()),
              "ash_super_absorbing_step")
        return out

    def apply_phase_screen(self, x, phase, amp=None, out=None):
        out = self._io(x, out)
        check(self.lib.ash_apply_phase_screen(self.h, ptr(x), ptr(out), ptr(phase), ptr(amp),
                                              _lib.current_stream()), "ash_apply_phase_screen")
        return out


def _check_field(t: torch.Tensor, shape) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError("expected a CUDA (HIP) torch tensor")
    if t.dtype != torch.float32 or tuple(t.shape) != tuple(shape) or not t.is_contiguous():
        raise ValueError(f"expected a contiguous float32 tensor of shape {tuple(shape)}, "
                         f"got {t.dtype} {tuple(t.shape)}")
