"""The reference's own validation grids on the GPU (general-size kernels).

The reference validates on odd and prime grids (validate_solver.py:80, :140;
validate_boundaries.py:98; validate_dc_invariance.py:27) that the FFT kernels do
not cover; these run the general-size kernels (direct DFTs) through the same C ABI:

* validate_dc_invariance.py:60-94, the reference's only pass/fail check: a uniform
  pulse on 32 x 32 x nT, nT in {120, 240, 480}, 64 full propagator steps; the DC bin
  must not grow and the peak |p| must agree across nT within 5 %;
* validate_solver.py tests 1 and 2 (51 x 51 x 101, 3 MHz, sequential and split
  linear water runs; Rusanov and KT nonlinear runs with attenPow = 1, the
  logarithmic Kramers-Kronig branch) against the FP64 oracle, rel-L2 <= 1e-4 on every
  output;
* validate_boundaries.py test 2 (81 x 81 x 101): the boundary/centre intensity
  ratios, published as 0.0183 / 0.0135 / 0.0094 / 0.0075;
* validate_solver.py test 8 (:557-695, nT = 801, N = 1): the global-max CFL
  sub-cycled KT step (_kt_flux_adaptive) at the test's step sizes against the
  oracle;
* the ASM half of validate_bowl.py (run_asm, :222-305): the TIPS bowl on the
  script's 365 x 365 x 425 grid, plane by plane and flat, against the FP64 torch twin.
"""
import numpy as np
import pytest

from oracle import asm_oracle as O
from tests.conftest import rel_l2
from tests.test_oracle_boundaries import REFLECTION_CASES, reflection_ratio, reflection_setup

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_dc_invariance_known_answer_on_gpu():
    import torch
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    peaks, ratios = [], []
    for nT in (120, 240, 480):
        nX = nY = 32
        f0, c0 = 1.0e6, 1500.0
        lam = c0 / f0
        dX = dY = lam / 8
        dZ = lam * 0.25
        dT = 1.0 / (f0 * 16)
        t = np.arange(nT) * dT
        env = np.exp(-((t - 3 / f0 * 0.5) / (3 / f0 / 4)) ** 2)
        pulse = (np.sin(2 * np.pi * f0 * t) * env).astype(np.float32)
        field = np.broadcast_to(pulse[None, None, :], (nX, nY, nT)).astype(np.float32).copy()
        p = SolverParams(dX=dX, dY=dY, dT=dT, c0=c0, useAdaptiveFiltering=False, useBoundaryLayer=False)
        plan = MarchPlan(field.shape, p)
        plan.set_dz(dZ)
        x, y = _dev(field), torch.empty((nX, nY, nT), dtype=torch.float32, device="cuda")
        for _ in range(64):
            plan.angular_spectrum_step(x, False, out=y)
            x, y = y, x
        got = x.cpu().numpy()
        src_dc = np.abs(np.fft.fft(field, axis=2)[..., 0]).max()
        peaks.append(float(np.max(np.abs(got))))
        ratios.append(float(np.abs(np.fft.fft(got, axis=2)[..., 0]).max()) / max(src_dc, 1e-30))
        # the same 64 steps in the FP64 oracle
        HH, _ = O.precalculate_mas(nX, nY, nT, dX, dY, dZ, dT, c0, adaptive_filtering=False)
        ref = field.astype(np.float64)
        for _ in range(64):
            ref = O.angular_spectrum_step(ref, HH, 1.0)
        e = rel_l2(got, ref)
        print(f"nT={nT}: peak {peaks[-1]:.5f}, DC ratio {ratios[-1]:.2e}, rel-L2 vs oracle {e:.2e}", flush=True)
        assert e < 1e-4
        plan.close()
    assert max(ratios) <= 1.0
    assert (max(peaks) - min(peaks)) / min(peaks) <= 0.05


def _validate_solver_ic(p0):
    """validate_solver.py:39-57 with its test 1/2 grid: 51 x 51 x 101, 3 MHz, dX = lambda/5."""
    f0, c0 = 3e6, 1500.0
    dX = c0 / f0 / 5
    dT = dX / (5 * c0)
    nX = nY = 51
    x = np.arange(nX) * dX - (nX - 1) * dX / 2
    t = np.arange(101) * dT
    t -= np.mean(t)
    w0 = 2 * np.pi * f0
    ap = ((np.abs(x)[:, None] <= 2.5e-3) & (np.abs(x)[None, :] <= 2.5e-3)).astype(float)
    burst = np.exp(-(1.05 * t * w0 / (1.5 * np.pi)) ** 4) * np.sin(t * w0) * p0
    field = (ap[:, :, None] * burst[None, None, :]).astype(np.float32)
    return field, dict(dX=dX, dY=dX, dT=dT, c0=c0, rho0=1000, f0=f0, boundaryFactor=0.15,
                       useAdaptiveFiltering=False, stabilityThreshold=0.2, stabilityRecoveryFactor=0.15,
                       dZmin=dX * 4)


CASES = [
    ("test1 sequential", 1e3, dict(beta=0.0, alpha0=-1, propDist=2e-2, useSplitStep=False)),
    ("test1 split", 1e3, dict(beta=0.0, alpha0=-1, propDist=2e-2, useSplitStep=True)),
    ("test2 rusanov", 0.03e6, dict(beta=3.5, alpha0=0.5, attenPow=1, propDist=3e-2, useSplitStep=True,
                                   fluxScheme='rusanov', useTVD=False)),
    ("test2 kt", 0.03e6, dict(beta=3.5, alpha0=0.5, attenPow=1, propDist=3e-2, useSplitStep=True,
                              fluxScheme='kt')),
]


@pytest.mark.parametrize("label,p0,kw", CASES, ids=[c[0] for c in CASES])
def test_validate_solver_grids_match_oracle(label, p0, kw):
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field, base = _validate_solver_ic(p0)
    got = angular_spectrum_solve(field, SolverParams(**base, **kw), verbose=False)
    ref = O.angular_spectrum_solve(field, O.OracleParams(**base, **kw), dtype=np.float64)
    np.testing.assert_allclose(got[5], ref[5], rtol=0, atol=1e-15)
    for i, n in enumerate(("field", "pnp", "ppp", "pI", "pIloss", None, "pax")):
        if n is None:
            continue
        e = rel_l2(got[i], ref[i])
        print(label, n, f"{e:.2e}", flush=True)
        assert e <= 1e-4, (n, e)


def test_boundary_reflection_ratios_on_gpu():
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field, base = reflection_setup()
    got = []
    for label, kw, published in REFLECTION_CASES:
        res = angular_spectrum_solve(field, SolverParams(**base, **kw), verbose=False)
        r = reflection_ratio(res[3][:, :, -1])
        print(f"{label}: {r:.4e} (published {published})", flush=True)
        assert len(res[5]) == 75
        assert abs(r - published) < 0.0051 * published + 5e-5, (label, r, published)
        got.append(r)
    assert got[0] > got[1] > got[2] > got[3]


@pytest.mark.parametrize("nsub_div", [16, 64, 256])
def test_validate_solver_test8_adaptive_kt(nsub_div):
    """validate_solver.py:557-695: u0 = 2.5 exp(-(t/0.18)^2) sin(8 pi t) on 801 samples of
    [-1, 1], N = 1, z = 0.05 in steps z/16 .. z/256 of _kt_flux_adaptive (global-max CFL
    sub-cycling: 7 sub-steps per step at z/16).  The reference's 1 x 1 field is held
    in a 2 x 2 plan (identical columns: the maximum is the same)."""
    import math
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    nT = 801
    t = np.linspace(-1.0, 1.0, nT)
    dT = t[1] - t[0]
    u0 = 2.5 * np.exp(-(t / 0.18) ** 2) * np.sin(2 * np.pi * 4 * t)
    u0[[0, -1]] = 0.0
    field = np.broadcast_to(u0.astype(np.float32), (2, 2, nT)).copy()
    # N = beta / (2 c0^3 rho0) = 1
    p = SolverParams(dX=1.0, dY=1.0, dT=dT, c0=1.0, rho0=0.5, beta=1.0, fluxScheme='kt')
    plan = MarchPlan(field.shape, p)
    dz = 0.05 / nsub_div
    x = _dev(field)
    ref = field.astype(np.float64)
    for _ in range(nsub_div):
        x = plan.nonlinear_step(x, _lib.NL_KT, dz, subcycle=_lib.SUBCYCLE_GLOBAL)
        ref = O.kt_flux_adaptive(ref, 1.0, dz, dT)
    got = x.cpu().numpy()
    nsub = math.ceil(2.5 * dz / dT / 0.5)
    e = rel_l2(got, ref)
    print(f"dZ = z/{nsub_div}: {nsub} sub-steps at the start; rel-L2 vs oracle {e:.2e}", flush=True)
    assert np.isfinite(got).all()
    assert e < 1e-5
    assert plan.subcycle_capped() == 0


# validate_bowl.py:40-65: the TIPS bowl transducer in water
BOWL = dict(f0=1e6, c0=1540.0, rho0=1000.0, roc=80e-3, outer=46e-3, inner=20.5e-3, n_el=8, ncycles=3, dur=2)


def _bowl_asm_setup(mode):
    """validate_bowl.py:222-290 (run_asm): the script's own grid (dX = lambda/5 over 112 mm,
    odd sizes; nT covering the pulse and the delay spread, odd, <= 601), dZ = dX over 80 mm,
    lossless water, Wendland + frequency-weighted + super-absorbing (0.5) boundaries."""
    from angularspectrum_heterogeneous_amd.sources import make_bowl_source, make_bowl_source_planes
    b = BOWL
    lam = b["c0"] / b["f0"]
    dX = lam / 5
    nX = int(np.ceil(112e-3 / dX))
    nX += 1 - nX % 2
    spread = (np.sqrt(b["roc"] ** 2 + b["outer"] ** 2) - b["roc"]) / b["c0"]
    dT = dX / (5 * b["c0"])
    nT = int(np.ceil((2 * b["ncycles"] / b["f0"] + spread + 3e-6) / dT))
    nT = min(nT + 1 - nT % 2, 601)
    x = np.arange(nX) * dX - (nX - 1) * dX / 2
    t = np.arange(nT) * dT - (nT - 1) * dT / 2
    z_inner = b["roc"] - np.sqrt(b["roc"] ** 2 - b["inner"] ** 2)
    focus = 50e-3 + z_inner
    p0 = 4e5 * 1.306
    common = dict(radius=b["outer"], roc=b["roc"], focus=focus, ncycles=b["ncycles"], dur=b["dur"],
                  inner_radius=b["inner"], n_elements=b["n_el"])
    if mode == "flat":
        field = make_bowl_source(x, x, t, b["f0"], b["c0"], p0, **common)
        planes = None
    else:
        sl, _ = make_bowl_source_planes(x, x, t, b["f0"], b["c0"], p0, dZ=dX, **common)
        field, planes = sl[0][1], sl[1:]
    kw = dict(dX=dX, dY=dX, dT=dT, c0=b["c0"], rho0=b["rho0"], beta=0.0, alpha0=-1.0, f0=b["f0"],
              propDist=80e-3, boundaryFactor=0.15, useSplitStep=True, useAdaptiveFiltering=True, useTVD=False,
              dZmin=dX, boundaryProfile='wendland', useFreqWeightedBoundary=True, useSuperAbsorbing=True,
              superAbsorbingStrength=0.5, sourcePlanes=planes)
    return field, kw, (nX, nT)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["planes", "flat"])
def test_validate_bowl_asm_scenario_against_fp64(mode):
    """The ASM half of validate_bowl.py (run_asm, :222-305) on its 365 x 365 x 425 grid
    (odd on every axis: the general-size kernels throughout), the bowl injected plane by
    plane (sparse mask + delay slices) or as the flat one-plane source, against the FP64
    torch twin of the oracle: rel-L2 <= 1e-4 on every output."""
    import time
    import torch
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    from oracle import torch_twin as TT
    field, kw, (nX, nT) = _bowl_asm_setup(mode)
    print(f"[bowl {mode}] grid {nX}x{nX}x{nT}, {len(kw['sourcePlanes'] or [])} source planes", flush=True)
    assert (nX, nT) == (365, 425)
    t0 = time.perf_counter()
    got = angular_spectrum_solve(field, SolverParams(**kw), verbose=False)
    t1 = time.perf_counter()
    ref = TT.angular_spectrum_solve(np.asarray(field), O.OracleParams(**kw), device="cuda", progress=100)
    torch.cuda.empty_cache()
    print(f"[bowl {mode}] gpu {t1 - t0:.1f} s ({len(got[5])} steps), fp64 twin {time.perf_counter() - t1:.1f} s",
          flush=True)
    np.testing.assert_allclose(got[5], ref[5], rtol=0, atol=1e-15)
    for i, n in enumerate(("field", "pnp", "ppp", "pI", "pIloss", None, "pax")):
        if n is None:
            continue
        e = rel_l2(got[i], ref[i])
        print(f"[bowl {mode}] {n} {e:.2e}", flush=True)
        assert e <= 1e-4, (n, e)
