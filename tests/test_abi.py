"""CPU checks of the C-ABI boundary: the library loads, exports every symbol the
header declares, and the ctypes structures match the C layout compiled by gcc.
No compute call is made (there is no GPU here); argument validation that
returns before touching the device is exercised."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aship.h")


def _header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ash_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from angularspectrum_heterogeneous_amd import _lib
    lib = _lib.load()
    declared = _header_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared


def test_abi_version_of_header_bindings_and_library_agree():
    """ash_abi_version is a host-only call: the loaded library, the header's
    ASH_ABI_VERSION and the bindings' ABI_VERSION must all be the same number."""
    from angularspectrum_heterogeneous_amd import _lib
    m = re.search(r"#define\s+ASH_ABI_VERSION\s+(\d+)", open(HEADER).read())
    assert m and int(m.group(1)) == _lib.ABI_VERSION
    assert _lib.load().ash_abi_version() == _lib.ABI_VERSION


def test_struct_layout_matches_c(tmp_path):
    from angularspectrum_heterogeneous_amd import _lib
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "aship.h"
#define P(s, f) printf(#s "." #f " %zu\n", offsetof(s, f));
int main(void) {
  printf("ash_desc %zu\n", sizeof(ash_desc));
  printf("ash_step_io %zu\n", sizeof(ash_step_io));
  P(ash_desc, nT) P(ash_desc, dX) P(ash_desc, f0) P(ash_desc, useSplitStep) P(ash_desc, freqFilterThreshold)
  P(ash_desc, useBoundaryLayer) P(ash_desc, superAbsorbingStrength) P(ash_desc, fluxScheme) P(ash_desc, cflTarget)
  P(ash_step_io, pnp_dev) P(ash_step_io, field_out_dev) P(ash_step_io, dz) P(ash_step_io, n_screens)
  P(ash_step_io, screen_phase_dev) P(ash_step_io, screen_amp_dev) P(ash_step_io, n_src_dense)
  P(ash_step_io, src_dense_dev) P(ash_step_io, src_mask_dev) P(ash_step_io, src_f0) P(ash_step_io, src_dur)
  return 0;
}
'''
    c = str(tmp_path / "abi_layout.c")
    exe = str(tmp_path / "abi_layout")
    open(c, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    assert int(got["ash_desc"]) == ctypes.sizeof(_lib.AshDesc)
    assert int(got["ash_step_io"]) == ctypes.sizeof(_lib.AshStepIO)
    for k, v in got.items():
        if "." in k:
            s, f = k.split(".")
            cls = _lib.AshDesc if s == "ash_desc" else _lib.AshStepIO
            assert getattr(cls, f).offset == int(v), k


def test_plan_create_rejects_unsupported_sizes_before_device_use():
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import desc_from_params
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    lib = _lib.load()
    p = SolverParams(dX=3e-4, dY=3e-4, dT=4e-8)
    h = ctypes.c_void_p()
    for shape in [(1, 51, 101), (64, 1025, 256), (64, 64, 3), (64, 64, 4096)]:
        d = desc_from_params(shape, p)
        assert lib.ash_plan_create(ctypes.byref(d), ctypes.byref(h)) == -2
        assert b"must be between" in lib.ash_last_error()


def test_error_codes_match_header():
    from angularspectrum_heterogeneous_amd import _lib
    txt = open(HEADER).read()
    for name in ("ASH_E_ARG", "ASH_E_UNSUPPORTED", "ASH_E_HIP", "ASH_E_STATE", "ASH_E_HOST"):
        m = re.search(r"#define\s+%s\s+\((-\d+)\)" % name, txt)
        assert m and int(m.group(1)) == getattr(_lib, name), name


def test_create_failing_on_the_device_returns_an_error_and_no_handle():
    """Without a device the first hipMalloc of a valid plan or network fails: the
    call returns ASH_E_HIP with the runtime's text, leaves the handle NULL and frees
    the partial object (no crash, nothing escapes the C ABI)."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: creation succeeds here")
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import desc_from_params
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    lib = _lib.load()
    d = desc_from_params((64, 64, 256), SolverParams(dX=3e-4, dY=3e-4, dT=4e-8))
    h = ctypes.c_void_p(1)
    assert lib.ash_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.ASH_E_HIP
    assert h.value is None and b"hipMalloc" in lib.ash_last_error()
    n = ctypes.c_void_p(1)
    assert lib.ash_net_create(256, 16, ctypes.byref(n)) == _lib.ASH_E_HIP
    assert n.value is None and b"hipMalloc" in lib.ash_last_error()


def test_argument_errors_are_reported_before_any_device_use():
    """Invalid arguments return ASH_E_ARG (or ASH_E_UNSUPPORTED) with a message,
    with or without a GPU: the checks run before any allocation or launch."""
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import desc_from_params
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    lib = _lib.load()
    h = ctypes.c_void_p()
    d = desc_from_params((64, 64, 256), SolverParams(dX=0.0, dY=3e-4, dT=4e-8))
    assert lib.ash_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.ASH_E_ARG
    assert lib.ash_plan_create(None, ctypes.byref(h)) == _lib.ASH_E_ARG
    assert lib.ash_source_field(None, None, None, 4, 8, 1e6, 1.0, 3.0, 2.0, None, None) == _lib.ASH_E_ARG
    assert b"null" in lib.ash_last_error()
    # CFL sub-cycling: KT only, known modes only
    d = desc_from_params((64, 64, 256), SolverParams(dX=3e-4, dY=3e-4, dT=4e-8, fluxScheme='kt'), "global")
    d.fluxScheme = 0
    assert lib.ash_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.ASH_E_ARG
    assert b"KT" in lib.ash_last_error()
    d.fluxScheme, d.subcycle = 1, 7
    assert lib.ash_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.ASH_E_ARG
    with pytest.raises(ValueError, match="KT"):
        desc_from_params((64, 64, 256), SolverParams(dX=3e-4, dY=3e-4, dT=4e-8), "column")
    with pytest.raises(ValueError, match="ash_subcycle"):
        desc_from_params((64, 64, 256), SolverParams(dX=3e-4, dY=3e-4, dT=4e-8, fluxScheme='kt'), "always")
    assert lib.ash_abi_version() == _lib.ABI_VERSION == 3


def test_solver_params_match_reference_fields_and_defaults():
    """Field names and defaults of SolverParams (ASS:35-76)."""
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    ref = dict(dX=0.0, dY=0.0, dT=0.0, c0=1500.0, rho0=1000.0, beta=3.5, alpha0=0.5, attenPow=1.0, f0=5e6,
               propDist=0.08, boundaryFactor=0.2, useSplitStep=True, useAdaptiveFiltering=True, useTVD=True,
               freqFilterThreshold=0.05, adaptiveFilterStrength=0.7, stabilityThreshold=0.2,
               stabilityRecoveryFactor=0.15, dZmin=1e-3, useBoundaryLayer=True, boundaryProfile='quadratic',
               useFreqWeightedBoundary=False, useSuperAbsorbing=False, superAbsorbingStrength=0.8,
               fluxScheme='rusanov', useObliquityCorrection=True, useNonlinearityObliquity=False,
               phaseScreens=None, sourcePlanes=None)
    p = SolverParams()
    for k, v in ref.items():
        assert getattr(p, k) == v, k
    extra = [f for f in p.__dataclass_fields__ if f not in ref]
    assert all(f.startswith("ash_") for f in extra)


def test_product_path_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "angularspectrum_heterogeneous_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", txt, re.M), f


def test_missing_library_fails_loudly(tmp_path):
    from angularspectrum_heterogeneous_amd import _lib
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _lib.load(str(tmp_path / "missing.so"))
    finally:
        _lib._lib = saved


def test_single_hip_runtime_in_process():
    """The library must share PyTorch's HIP runtime (it is handed torch's
    streams and device pointers): after loading, one libamdhip64 is mapped."""
    pytest.importorskip("torch")
    from angularspectrum_heterogeneous_amd import _lib
    _lib.load()
    maps = open("/proc/self/maps").read()
    libs = {l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l}
    assert len(libs) == 1, libs


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before any device use."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_wavenumber_grids_equal_numpys_bit_for_bit():
    """The plan's kx/ky tables (host FP64) equal the reference's NumPy grids
    (ASS:405-408: linspace / (n-1) / d * 2 pi minus np.mean) to the last bit, in
    natural FFT order; a grid value one ulp off can flip a mode at the evanescent
    cutoff (k_z = 0) of the attenuation's obliquity factor."""
    import numpy as np
    from angularspectrum_heterogeneous_amd import _lib
    lib = _lib.load()
    for d in (3.08e-4, 1e-4, 0.1 * 1.5e-3 / 3, 1.0):
        for n in list(range(2, 140)) + [192, 255, 256, 257, 365, 384, 511, 601, 1000, 1024]:
            k = np.linspace(0, n - 1, n) / (n - 1) / d * 2 * np.pi
            k -= np.mean(k)
            want = k[(np.arange(n) + n // 2) % n]          # natural bin i <- fftshifted index (i + n//2) % n
            got = (ctypes.c_double * n)()
            assert lib.ash_centred_wavenumbers(n, d, got) == 0
            np.testing.assert_array_equal(np.frombuffer(got, dtype=np.float64), want, err_msg=f"n={n} d={d}")
