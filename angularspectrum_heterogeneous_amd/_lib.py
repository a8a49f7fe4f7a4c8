"""ctypes binding of the gfx950 HIP library ``libaship.so`` (C ABI in include/aship.h).

There is no fallback: if the library is missing or fails to load, every entry
point raises.  Device pointers are passed as integers (``tensor.data_ptr()``)
and the HIP stream as the raw ``hipStream_t`` of PyTorch's current stream.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaship.so")

ASH_OK = 0
ASH_E_ARG, ASH_E_UNSUPPORTED, ASH_E_HIP, ASH_E_STATE, ASH_E_HOST = -1, -2, -3, -4, -5
NL_NONE, NL_KT, NL_RUSANOV, NL_TVD, NL_KT_MINMOD = 0, 1, 2, 3, 4
SUBCYCLE_OFF, SUBCYCLE_COLUMN, SUBCYCLE_GLOBAL = 0, 1, 2
TOF_OFF, TOF_ENVELOPE, TOF_MATCHED = 0, 1, 2
MAX_SUBSTEPS = 1024
PROFILE_QUADRATIC, PROFILE_WENDLAND = 0, 1

c_i32 = ctypes.c_int32
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class AshDesc(ctypes.Structure):
    _fields_ = [
        ("nX", c_i32), ("nY", c_i32), ("nT", c_i32),
        ("dX", c_dbl), ("dY", c_dbl), ("dT", c_dbl),
        ("c0", c_dbl), ("rho0", c_dbl), ("beta", c_dbl),
        ("alpha0", c_dbl), ("attenPow", c_dbl), ("f0", c_dbl),
        ("boundaryFactor", c_dbl),
        ("useSplitStep", c_i32), ("useAdaptiveFiltering", c_i32), ("useTVD", c_i32),
        ("freqFilterThreshold", c_dbl), ("adaptiveFilterStrength", c_dbl),
        ("useBoundaryLayer", c_i32), ("boundaryProfile", c_i32),
        ("useFreqWeightedBoundary", c_i32), ("useSuperAbsorbing", c_i32),
        ("superAbsorbingStrength", c_dbl),
        ("fluxScheme", c_i32), ("useObliquityCorrection", c_i32),
        ("subcycle", c_i32), ("cflTarget", c_dbl),
        ("useNonlinearityObliquity", c_i32),
    ]


class AshStepIO(ctypes.Structure):
    _fields_ = [
        ("row", c_i32), ("max_rows", c_i32),
        ("pnp_dev", c_vp), ("ppp_dev", c_vp), ("pI_dev", c_vp), ("pIloss_dev", c_vp),
        ("pax_dev", c_vp), ("stats_dev", c_vp), ("field_out_dev", c_vp),
        ("dz", c_dbl),
        ("n_screens", c_i32),
        ("screen_phase_dev", c_vp * 4), ("screen_amp_dev", c_vp * 4),
        ("n_src_dense", c_i32),
        ("src_dense_dev", c_vp * 4),
        ("src_mask_dev", c_vp), ("src_delay_dev", c_vp), ("taxis_dev", c_vp),
        ("src_f0", c_dbl), ("src_p0", c_dbl), ("src_ncycles", c_dbl), ("src_dur", c_dbl),
        ("tof_dev", c_vp),
    ]


# (name, argtypes) of every exported symbol; restype is c_int32 unless noted
_PROTOS = {
    "ash_abi_version": ([], c_i32),
    "ash_last_error": ([], ctypes.c_char_p),
    "ash_centred_wavenumbers": ([c_i32, c_dbl, ctypes.POINTER(c_dbl)], c_i32),
    "ash_plan_create": ([ctypes.POINTER(AshDesc), ctypes.POINTER(c_vp)], c_i32),
    "ash_plan_destroy": ([c_vp], None),
    "ash_plan_set_dz": ([c_vp, c_dbl], c_i32),
    "ash_load_field": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_store_field": ([c_vp, c_vp, c_vp], c_i32),
    "ash_step": ([c_vp, ctypes.POINTER(AshStepIO), c_vp], c_i32),
    "ash_march_step": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_attenuation_step": ([c_vp, c_vp, c_vp, c_dbl, c_vp], c_i32),
    "ash_angular_spectrum_step": ([c_vp, c_vp, c_vp, c_i32, c_vp], c_i32),
    "ash_nonlinear_step": ([c_vp, c_vp, c_vp, c_i32, c_dbl, c_dbl, c_i32, c_vp], c_i32),
    "ash_max_abs": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_plan_subcycle_capped": ([c_vp, ctypes.POINTER(ctypes.c_int64)], c_i32),
    "ash_nonlinear_step_noalloc": ([ctypes.c_int64, c_i32, c_dbl, c_dbl, c_vp, c_vp, c_i32, c_dbl, c_dbl, c_i32,
                                    c_dbl, c_vp, c_vp], c_i32),
    "ash_freq_weighted_boundary_step": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_super_absorbing_step": ([c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_apply_phase_screen": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_i32),
    "ash_source_field": ([c_vp, c_vp, c_vp, ctypes.c_int64, c_i32, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp], c_i32),
    "ash_profile_step": ([c_vp, ctypes.POINTER(AshStepIO), c_i32, ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(c_i32), c_vp], c_i32),
    "ash_plan_workspace_bytes": ([c_vp], ctypes.c_int64),
    "ash_plan_autotune": ([c_vp, c_vp, c_vp], c_i32),
    "ash_plan_set_launch": ([c_vp, c_vp], c_i32),
    "ash_batch_create": ([ctypes.POINTER(c_vp)], c_i32),
    "ash_batch_destroy": ([c_vp], None),
    "ash_batch_step": ([c_vp, c_vp, c_vp, c_i32, c_vp], c_i32),
    "ash_batch_set_branches": ([c_vp, c_i32], c_i32),
    "ash_net_param_count": ([], ctypes.c_int64),
    "ash_net_create": ([c_i32, c_i32, ctypes.POINTER(c_vp)], c_i32),
    "ash_net_destroy": ([c_vp], None),
    "ash_net_set_params": ([c_vp, c_vp, c_vp], c_i32),
    "ash_net_get_params": ([c_vp, c_vp, c_vp], c_i32),
    "ash_net_forward": ([c_vp, c_vp, c_i32, c_vp, c_vp, c_vp], c_i32),
    "ash_net_loss_grad": ([c_vp, c_vp, c_vp, c_i32, c_dbl, c_vp, c_vp, c_vp], c_i32),
    "ash_net_adam": ([c_vp, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, c_i32, c_vp],
                     c_i32),
    "ash_net_apply": ([c_vp, c_vp, c_i32, c_dbl, c_dbl, c_dbl, c_vp, c_vp], c_i32),
    "ash_plan_set_learned": ([c_vp, c_vp, c_dbl], c_i32),
    "ash_plan_set_tof": ([c_vp, c_i32, c_dbl, c_dbl, c_dbl, c_vp, c_dbl, c_vp], c_i32),
}

KERNEL_KINDS = ("xfft_fwd", "ypass", "xfft_inv", "ysa", "column_nl", "column_taper", "column_final",
                "column_other", "net")

EXPORTED = tuple(_PROTOS)

# ASH_ABI_VERSION of include/aship.h that the prototypes and structures above follow
ABI_VERSION = 3

_lib: Optional[ctypes.CDLL] = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the library once; raises if it is absent (there is no CPU path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with angularspectrum_heterogeneous_amd._build.build() "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    # One HIP runtime per process: PyTorch's ROCm wheel bundles libamdhip64
    # (SONAME libamdhip64.so.7).  Importing torch first registers that runtime,
    # and this library's libamdhip64.so.7 dependency then binds to it; loaded
    # the other way round, a second runtime would be mapped and the caller's
    # hipStream_t handles would belong to a different runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (args, res) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    got = lib.ash_abi_version()
    if got != ABI_VERSION:
        raise RuntimeError(f"{path} implements ABI version {got}, the bindings expect {ABI_VERSION}: "
                           "rebuild it with angularspectrum_heterogeneous_amd._build.build()")
    _lib = lib
    return lib


class AshError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != ASH_OK:
        msg = load().ash_last_error()
        raise AshError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(t) -> Optional[int]:
    """Device pointer of a torch tensor (None passes NULL)."""
    return None if t is None else t.data_ptr()


def current_stream() -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream
