"""The entry point that allocates nothing, ash_nonlinear_step_noalloc, against the
existing entry point it varies, ash_nonlinear_step, through the C ABI: two calls in
a row on one GPU (the second on a smaller input, so the fixed scratch holds the first
call's contents), in each sub-cycling mode.  The two entry points run the same
kernels, so their outputs agree bit for bit; the reported maximum equals
ash_max_abs and the sub-step counter is cleared by every call."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(shape, p0, seed):
    rng = np.random.default_rng(seed)
    nX, nY, nT = shape
    t = (np.arange(nT) - (nT - 1) / 2) * 4e-8
    burst = np.exp(-(t / (40 * 4e-8)) ** 2) * np.sin(2 * np.pi * 1e6 * t)
    amp = p0 * (0.2 + rng.random((nX, nY)))
    f = amp[:, :, None] * burst[None, None, :] + 1e-3 * p0 * rng.standard_normal(shape)
    return f.astype(np.float32)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_noalloc_matches_plan_entry_point_on_two_calls(mode):
    import torch
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.workloads import params_kwargs
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    lib = _lib.load()
    kw = params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=0.02, dZmin=3.08e-4)
    p = SolverParams(**kw)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    dz = 1e-2
    stream = _lib.current_stream()
    for shape, p0, seed in [((128, 128, 1024), 2e7, 0), ((64, 64, 256), 3e7, 1)]:
        print(f"mode {mode} shape {shape}", flush=True)
        u = torch.from_numpy(_field(shape, p0, seed)).cuda()
        plan = MarchPlan(shape, p)
        want = plan.nonlinear_step(u, _lib.NL_KT, dz, subcycle=mode)
        got = torch.empty_like(u)
        rep = torch.full((2,), 12345, dtype=torch.int32, device="cuda")
        ncols = shape[0] * shape[1]
        _lib.check(lib.ash_nonlinear_step_noalloc(ncols, shape[2], N, p.dT, u.data_ptr(), got.data_ptr(),
                                                  _lib.NL_KT, dz, 1.0, mode, 0.5, rep.data_ptr(), stream),
                   "ash_nonlinear_step_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        r = rep.cpu().numpy()
        gmax = float(np.frombuffer(r[:1].tobytes(), dtype=np.float32)[0])
        assert gmax == (plan.max_abs(u) if mode == 2 else 0.0)
        assert r[1] == 0
        if mode == 2:
            assert N * gmax * dz / p.dT > 0.5       # the sub-cycling engaged
        plan.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from angularspectrum_heterogeneous_amd import _lib
    lib = _lib.load()
    rc = lib.ash_nonlinear_step_noalloc(64 * 16384 + 1, 256, 1e-12, 4e-8, 1, 1, _lib.NL_KT, 1e-3, 1.0, 2, 0.5,
                                        None, None)
    assert rc == _lib.ASH_E_UNSUPPORTED and b"1048576" in lib.ash_last_error()
    rc = lib.ash_nonlinear_step_noalloc(64, 300, 1e-12, 4e-8, 1, 1, _lib.NL_KT, 1e-3, 1.0, 2, 0.5, None, None)
    assert rc == _lib.ASH_E_UNSUPPORTED
    rc = lib.ash_nonlinear_step_noalloc(64, 256, 1e-12, 4e-8, 1, 1, _lib.NL_RUSANOV, 1e-3, 1.0, 2, 0.5, None, None)
    assert rc == _lib.ASH_E_ARG
