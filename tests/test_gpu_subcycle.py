"""GPU parity of CFL sub-cycling of the KT step.

* Global mode (``ASH_SUBCYCLE_GLOBAL``, ``ash_subcycle='global'``): the reference's
  ``_kt_flux_adaptive`` (ASS:842-859), one sub-step count for the whole field from
  its max|u|, reduced on the device.  Its only caller in the reference is
  validate_solver.py test 8 (:557-695); its oracle is ``kt_flux_adaptive``.
* Per-column mode (``ash_subcycle='column'``): this library's extension; its oracle
  is ``kt_flux_subcycled_columns`` (equal to ``kt_flux`` wherever a column's CFL
  number is at or below the target, pinned on the CPU).
* The kernels' bound of 1024 sub-steps per dZ is counted on the device and raised
  by the driver (``SubcycleLimitError``); the reference's host loop has no bound.

Bounds: rel-L2 1e-5 on one sub-cycled operator, 1e-4 per output on a short solve.
"""
import math

import numpy as np
import pytest

from oracle import asm_oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _beam(shape, p0, seed=0):
    from angularspectrum_heterogeneous_amd.workloads import axes, tone_burst
    nX, nY, nT = shape
    rng = np.random.default_rng(seed)
    x, y, t = axes(nX, 3.08e-4), axes(nY, 3.08e-4), axes(nT, 4e-8)
    ap = np.exp(-(x[:, None] ** 2 + y[None, :] ** 2) / (2 * (0.15 * nX * 3.08e-4) ** 2))
    f = ap[:, :, None] * tone_burst(t, 1e6, p0, 3)[None, None, :]
    f += 0.001 * p0 * rng.standard_normal(shape)
    return f.astype(np.float32)


def test_subcycled_kt_step_matches_oracle():
    import torch
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.workloads import params_kwargs
    shape = (64, 64, 256)
    p = O.OracleParams(**params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=0.02, dZmin=3.08e-4))
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    dz = 1e-2
    u = _beam(shape, 3e7)
    nsub = np.ceil(N * np.abs(u.astype(np.float64)).max(axis=2) * dz / p.dT / 0.5)
    print("sub-steps per column: min", int(nsub.min()), "max", int(nsub.max()), flush=True)
    assert nsub.min() <= 1 and 4 <= nsub.max() <= 64      # columns differ, all well below the cap
    plan = MarchPlan(shape, p)
    plan.set_dz(dz)
    got = plan.nonlinear_step(torch.from_numpy(u).cuda(), _lib.NL_KT, dz, subcycle=1).cpu().numpy()
    ref = O.kt_flux_subcycled_columns(u.astype(np.float64), N, dz, p.dT, cfl_target=0.5)
    plain = O.kt_flux(u.astype(np.float64), N, dz, p.dT)
    e, e_plain = rel_l2(got, ref), rel_l2(plain, ref)
    print(f"rel-L2 vs sub-cycled oracle {e:.2e}; plain KT differs by {e_plain:.2e}", flush=True)
    assert e < 1e-5
    assert e_plain > 10 * e                                # the sub-cycling is what is measured


@pytest.mark.parametrize("beta", [800.0])
def test_subcycled_solve_matches_oracle(beta):
    """A short focused march whose nonlinearity puts the beam's columns above the CFL
    target: beta raised so that N max|p| dZ / dT ~ 0.8, and stabilityThreshold raised
    above it (the driver otherwise keeps every step's CFL number below 0.2, ASS:1418-1427,
    where sub-cycling at a 0.5 target never engages)."""
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field = W.focused_piston(nX=64, nY=64, nT=256, radius=8e-3, focus=15e-3, p0=1e6, ncycles=3)
    kw = W.params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=1.0e-3, dZmin=3.08e-4, beta=beta,
                         fluxScheme='kt', stabilityThreshold=2.0)
    N = beta / (2 * kw.get('c0', 1540.0) ** 3 * kw.get('rho0', 1000.0))
    cfl = N * float(np.abs(field).max()) * 3.08e-4 / 4e-8
    print(f"peak CFL {cfl:.2f}", flush=True)
    assert cfl > 0.5 and math.ceil(cfl / 0.5) < 1024
    got = angular_spectrum_solve(field, SolverParams(ash_subcycle='column', **kw), verbose=False)
    ref = O.angular_spectrum_solve(np.asarray(field), O.OracleParams(**kw), dtype=np.float64,
                                   kt_subcycle='column')
    np.testing.assert_allclose(got[5], ref[5], rtol=0, atol=1e-15)
    for i, n in enumerate(("field", "pnp", "ppp", "pI", "pIloss")):
        e = rel_l2(got[i], ref[i])
        print(n, f"{e:.2e}", flush=True)
        assert e <= 1e-4, (n, e)


def _plan(shape, **kw):
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.workloads import params_kwargs
    p = O.OracleParams(**params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=0.02, dZmin=3.08e-4, **kw))
    return MarchPlan(shape, p), p


@pytest.mark.parametrize("shape,p0,dz", [((64, 64, 256), 3e7, 1e-2), ((64, 64, 1024), 2e7, 1e-2),
                                         ((128, 64, 512), 2e7, 4e-3)])
def test_global_subcycled_kt_step_matches_kt_flux_adaptive(shape, p0, dz):
    import torch
    from angularspectrum_heterogeneous_amd import _lib
    plan, p = _plan(shape)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    u = _beam(shape, p0, seed=1)
    cfl = N * float(np.abs(u).max()) * dz / p.dT
    nsub = math.ceil(cfl / 0.5)
    print(f"{shape}: global CFL {cfl:.2f} -> {nsub} sub-steps", flush=True)
    assert 2 <= nsub <= 64
    x = torch.from_numpy(u).cuda()
    got = plan.nonlinear_step(x, _lib.NL_KT, dz, subcycle=_lib.SUBCYCLE_GLOBAL).cpu().numpy()
    ref = O.kt_flux_adaptive(u.astype(np.float64), N, dz, p.dT, cfl_target=0.5)
    one = O.kt_flux(u.astype(np.float64), N, dz, p.dT)
    cols = O.kt_flux_subcycled_columns(u.astype(np.float64), N, dz, p.dT, cfl_target=0.5)
    e = rel_l2(got, ref)
    print(f"rel-L2 vs kt_flux_adaptive {e:.2e}; one step differs by {rel_l2(one, ref):.2e}, "
          f"per-column by {rel_l2(cols, ref):.2e}", flush=True)
    assert e < 1e-5
    assert rel_l2(one, ref) > 10 * e and rel_l2(cols, ref) > 10 * e
    assert plan.subcycle_capped() == 0
    # ash_max_abs reads the field without touching the plan state
    assert plan.max_abs(x) == float(np.abs(u).max())


def test_global_subcycled_solve_matches_oracle():
    """The march with ``ash_subcycle='global'``: the maximum is taken on the field the
    nonlinear pass sees (after the first half step and the taper), as
    ``march_step_split_kt`` with ``_kt_flux_adaptive`` in place of ``_kt_flux``."""
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field = W.focused_piston(nX=64, nY=64, nT=256, radius=8e-3, focus=15e-3, p0=1e6, ncycles=3)
    kw = W.params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=1.0e-3, dZmin=3.08e-4, beta=800.0,
                         fluxScheme='kt', stabilityThreshold=2.0)
    got = angular_spectrum_solve(field, SolverParams(ash_subcycle='global', **kw), verbose=False)
    ref = O.angular_spectrum_solve(np.asarray(field), O.OracleParams(**kw), dtype=np.float64,
                                   kt_subcycle='global')
    col = O.angular_spectrum_solve(np.asarray(field), O.OracleParams(**kw), dtype=np.float64,
                                   kt_subcycle='column')
    for i, n in enumerate(("field", "pnp", "ppp", "pI", "pIloss")):
        e = rel_l2(got[i], ref[i])
        print(n, f"{e:.2e} (per-column oracle differs by {rel_l2(col[i], ref[i]):.2e})", flush=True)
        assert e <= 1e-4, (n, e)
    assert rel_l2(col[0], ref[0]) > 100 * rel_l2(got[0], ref[0])   # the two modes are told apart


@pytest.mark.parametrize("mode", ["column", "global"])
def test_subcycle_bound_is_counted_and_raised(mode):
    """A column whose CFL number asks for more than 1024 sub-steps takes 1024 and is
    counted on the device; the operator reports the count, the driver raises."""
    import torch
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.solver import SubcycleLimitError
    shape = (64, 64, 256)
    plan, p = _plan(shape)
    u = _beam(shape, 1e5, seed=2)
    u[32, 32] *= 1e7                  # one column far beyond the bound
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    q = N * np.abs(u.astype(np.float64)).max(axis=2) * 1e-2 / p.dT / 0.5
    assert (q > 1024).sum() == 1 and q.max() > 1e4
    x = torch.from_numpy(u).cuda()
    sub = _lib.SUBCYCLE_COLUMN if mode == "column" else _lib.SUBCYCLE_GLOBAL
    plan.nonlinear_step(x, _lib.NL_KT, 1e-2, subcycle=sub)
    assert plan.subcycle_capped() == 1
    plan.nonlinear_step(x, _lib.NL_KT, 1e-2, subcycle=sub)
    assert plan.subcycle_capped() == 2
    plan.load(torch.zeros_like(x))   # a new march clears the count
    assert plan.subcycle_capped() == 0
    with pytest.raises(ValueError, match="KT"):
        plan.nonlinear_step(x, _lib.NL_RUSANOV, 1e-2, subcycle=sub)
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field = W.focused_piston(nX=64, nY=64, nT=256, radius=8e-3, focus=15e-3, p0=1e6, ncycles=3)
    kw = W.params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=1.0e-3, dZmin=3.08e-4, beta=1e6,
                         fluxScheme='kt', stabilityThreshold=1e9)
    with pytest.raises(SubcycleLimitError):
        angular_spectrum_solve(field, SolverParams(ash_subcycle=mode, **kw), verbose=False)
