"""Batched solves of uneven transmit configurations vs single-instance solves.

Each batched instance runs the same kernels, in the same order, as its
single-instance ``angular_spectrum_solve``; only the launch (one graph for the
whole batch per z-step) differs, so the outputs must agree to the last bit
(rel-L2 0, asserted at <= 1e-7).  With the autotuner a plan may take the other
radix order of a transform, which rounds differently in float32: <= 2e-5.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

NAMES = ("field", "pnp", "ppp", "pI", "pIloss", "zaxis", "pax")


def _compare(instances, results, tol=1e-7):
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    for k, ((field, kw), got) in enumerate(zip(instances, results)):
        ref = angular_spectrum_solve(field, SolverParams(**kw), verbose=False)
        for name, a, b in zip(NAMES, got, ref):
            assert a.shape == b.shape, (k, name)
            assert rel_l2(a, b) <= tol, (k, name)


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_matches_single_solves(autotune):
    print(f"[batch] uneven instances, autotune={autotune}", flush=True)
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.batch import batch_solve
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    spec = W.config4(n=5, seed=7, propDist=3e-3, sizes=(64, 128, 192), lengths=(256, 512))
    # uneven propagation distances too: instances leave the batch at different steps
    instances = [(f, dict(kw, propDist=kw["propDist"] * (1 + 0.4 * k))) for k, (f, kw, _) in enumerate(spec)]
    results = batch_solve([(f, SolverParams(**kw)) for f, kw in instances], autotune=autotune)
    _compare(instances, results, 2e-5 if autotune else 1e-7)


def test_batch_with_screen_step():
    """An instance whose step crosses a phase screen is enqueued directly (no graph) that
    step, while the others replay; all still equal their single solves."""
    print("[batch] screen step", flush=True)
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.batch import batch_solve
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    f1, kw1 = W.config1(propDist=4e-3)
    rng = np.random.default_rng(0)
    ph = rng.uniform(-1, 1, size=f1.shape[:2]).astype(np.float32)
    kw1s = dict(kw1, phaseScreens=[(1.5e-3, ph, None)])
    f2, kw2 = W.config1(propDist=3e-3)
    instances = [(f1, kw1s), (f2 * 2.0, kw2), (f1, kw1)]
    results = batch_solve([(f, SolverParams(**kw)) for f, kw in instances])
    _compare(instances, results)
