"""GPU parity of per-column CFL sub-cycling of the KT step (``ash_subcycle='column'``).

Not collected by pytest (the file name does not start with ``test_``): the mode has
not run on a GPU yet, so this file is run by name, alone, before it joins the suite:

    python -u -m pytest tests/wip_gpu_subcycle.py -x -v --timeout 120 --timeout-method thread

The reference defines the global-max variant (``_kt_flux_adaptive``, ASS:842-859) but its
driver never calls it; the per-column form is this library's opt-in extension and its
oracle is ``kt_flux_subcycled_columns`` (equal to ``kt_flux`` wherever a column's CFL
number is at or below the target, pinned on the CPU).  Bounds: rel-L2 1e-5 on one
sub-cycled operator, 1e-4 per output on a short solve.
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
