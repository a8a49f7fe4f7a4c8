"""End-to-end parity of ``angular_spectrum_solve`` (GPU) with the FP64 oracle driver.

The oracle restates the reference driver (ASS:1217-1567) in NumPy float64.
Bound: relative L2 error <= 1e-4 separately for field, pnp, ppp, pI, pIloss
and pax (the single-precision march against the FP64 oracle); zaxis equal.
"""
import numpy as np
import pytest

from oracle import asm_oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4
NAMES = ("field", "pnp", "ppp", "pI", "pIloss", "zaxis", "pax")


def _run_both(field, kw, tol=TOL, **solve_kw):
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    got = angular_spectrum_solve(field, SolverParams(**kw), verbose=False, **solve_kw)
    ref = O.angular_spectrum_solve(np.asarray(field), O.OracleParams(**kw), dtype=np.float64)
    assert got[5].shape == ref[5].shape, (got[5].shape, ref[5].shape)
    np.testing.assert_allclose(got[5], ref[5], rtol=0, atol=1e-15)
    errs = {}
    for i, n in enumerate(NAMES):
        if n == "zaxis":
            continue
        assert got[i].shape == ref[i].shape, (n, got[i].shape, ref[i].shape)
        errs[n] = rel_l2(got[i], ref[i])
    print({k: f"{v:.2e}" for k, v in errs.items()})
    for n, e in errs.items():
        assert e <= tol, (n, e)
    return got, ref, errs


def test_config1_full_run():
    from angularspectrum_heterogeneous_amd import workloads as W
    field, kw = W.config1()
    _run_both(field, kw)


def test_all_epilogue_features_partial_step():
    from angularspectrum_heterogeneous_amd import workloads as W
    field, kw = W.config1()
    nX, nY, nT = field.shape
    rng = np.random.default_rng(7)
    phase = rng.uniform(-2, 2, (nX, nY)).astype(np.float32)
    amp = rng.uniform(0.3, 1.0, (nX, nY)).astype(np.float32)
    src = (0.3 * W.gaussian_beam(p0=1e5)).astype(np.float32)
    kw.update(propDist=0.0101, boundaryProfile='quadratic', useFreqWeightedBoundary=True,
              useSuperAbsorbing=True, phaseScreens=[(2.2e-3, phase, amp), (2.3e-3, -phase)],
              sourcePlanes=[(1.5e-3, src)], fluxScheme='rusanov', useTVD=True)
    _run_both(field, kw)


def test_sequential_water():
    from angularspectrum_heterogeneous_amd import workloads as W
    field, kw = W.config1()
    kw.update(useSplitStep=False, alpha0=-1.0, propDist=0.006, useAdaptiveFiltering=False)
    _run_both(field, kw)


def test_split_standard_rusanov_no_obliquity():
    from angularspectrum_heterogeneous_amd import workloads as W
    field, kw = W.config1()
    kw.update(fluxScheme='rusanov', useTVD=False, useObliquityCorrection=False, attenPow=1.0,
              propDist=0.006, useBoundaryLayer=False)
    _run_both(field, kw)


def test_sparse_bowl_sources_match_dense():
    """Sparse (mask + delay) slice injection equals injecting the dense slices."""
    from angularspectrum_heterogeneous_amd import make_bowl_source_planes
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    nX = nY = 64
    nT = 256
    dX, dT = 3.08e-4, 4e-8
    x, y, t = W.axes(nX, dX), W.axes(nY, dX), W.axes(nT, dT)
    planes, depth = make_bowl_source_planes(x, y, t, 1e6, 1540.0, 3e5, 9e-3, 20e-3, dX, ncycles=3,
                                            inner_radius=2e-3, focus=15e-3, n_elements=3)
    assert len(planes) > 3
    kw = W.params_kwargs(dX=dX, dY=dX, dT=dT, propDist=depth + 2 * dX, dZmin=dX)
    init = np.asarray(planes[0][1])
    sparse = angular_spectrum_solve(init, SolverParams(sourcePlanes=planes[1:], **kw), verbose=False)
    dense_planes = [(z, np.asarray(s)) for z, s in planes[1:]]
    dense = angular_spectrum_solve(init, SolverParams(sourcePlanes=dense_planes, **kw), verbose=False)
    for i, n in enumerate(NAMES):
        e = rel_l2(sparse[i], dense[i])
        print(n, e)
        assert e < 1e-6, (n, e)
    # and the dense path against the oracle
    kw2 = dict(kw, sourcePlanes=dense_planes)
    _run_both(init, kw2)


def test_stability_error_raised():
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, StabilityError, angular_spectrum_solve
    field, kw = W.config1()
    kw.update(stabilityThreshold=1e-6, propDist=0.003)
    with pytest.raises(StabilityError):
        angular_spectrum_solve(field, SolverParams(**kw), verbose=False)
    with pytest.raises(O.StabilityError):
        O.angular_spectrum_solve(field, O.OracleParams(**kw))


def test_callback_sees_each_step():
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field, kw = W.config1()
    kw.update(propDist=0.002)
    seen = []

    def cb(cc, z, f):
        seen.append((cc, z, float(f.abs().max().item())))

    out = angular_spectrum_solve(field, SolverParams(**kw), verbose=False, per_step_callback=cb)
    assert [s[0] for s in seen] == list(range(len(out[5])))
    np.testing.assert_allclose([s[1] for s in seen], out[5], rtol=0, atol=1e-15)
    assert abs(seen[-1][2] - np.abs(out[0]).max()) <= 1e-6 * np.abs(out[0]).max()


def test_tof_arguments_rejected():
    """A reference trace needs its peak time (ASS:1392-1395); the trace and the
    time axis must have nT samples and the axis must be uniform."""
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field, kw = W.config1()
    nT = field.shape[2]
    with pytest.raises(ValueError, match="tof_t_ref_peak_s"):
        angular_spectrum_solve(field, SolverParams(**kw), verbose=False, tof_ref_trace=np.zeros(nT))
    with pytest.raises(ValueError, match="tof_ref_trace"):
        angular_spectrum_solve(field, SolverParams(**kw), verbose=False, tof_ref_trace=np.zeros(nT - 1),
                               tof_t_ref_peak_s=1e-6)
    with pytest.raises(ValueError, match="taxis"):
        angular_spectrum_solve(field, SolverParams(**kw), verbose=False, tof_env_ratio=0.5,
                               taxis=np.arange(nT + 1) * kw["dT"])
    warped = np.arange(nT) * kw["dT"]
    warped[nT // 2:] += 0.5 * kw["dT"]
    with pytest.raises(ValueError, match="uniform"):
        angular_spectrum_solve(field, SolverParams(**kw), verbose=False, tof_env_ratio=0.5, taxis=warped)


def _focused_small(p0=1e6):
    from angularspectrum_heterogeneous_amd import workloads as W
    field = W.focused_piston(nX=64, nY=64, nT=256, radius=8e-3, focus=15e-3, p0=p0, ncycles=3)
    kw = W.params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=0.008, dZmin=3.08e-4)
    return field, kw


@pytest.mark.parametrize("variant", ["split_kt", "split_tvd", "sequential"])
def test_nonlinearity_obliquity(variant):
    """useNonlinearityObliquity (ASS:923-965, :1374-1377): N scaled by the beam's power-
    weighted <k/k_z> before every nonlinear step, on a strongly focused beam."""
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    print(f"[solve] nonlinear obliquity {variant}", flush=True)
    field, kw = _focused_small()
    if variant == "split_tvd":
        kw.update(fluxScheme='rusanov', useTVD=True)
    elif variant == "sequential":
        kw.update(useSplitStep=False, fluxScheme='rusanov')
    kw_obl = dict(kw, useNonlinearityObliquity=True)
    got, _, _ = _run_both(field, kw_obl)
    plain = angular_spectrum_solve(field, SolverParams(**kw), verbose=False)
    # the scale matters here: the two runs differ far beyond the parity bound
    assert rel_l2(got[0], plain[0]) > 10 * TOL
