"""Operator-level parity of the gfx950 kernels against the FP64 oracle.

Each entry point of the C ABI is called on a seeded field and compared with
the oracle's restatement of the reference function it replaces.  The bound is
relative L2 error; single-precision kernels on one operator land at ~1e-7, the
tolerance is 2e-6 (5e-6 for the multi-sweep composite step).  GEN sizes run the
general-size kernels on one, two or all three axes (odd, prime and 2^a 3^b 5^c
lengths, odd nT without a Nyquist bin); their direct DFTs accumulate n terms per
output in float32, so their bounds are 5e-6 (1e-5 composite).
"""
import numpy as np
import pytest

from oracle import asm_oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [(64, 64, 256), (128, 192, 512), (256, 256, 1024), (384, 64, 2048), (64, 384, 512)]
GEN = [(51, 51, 101), (32, 32, 120), (64, 51, 256), (51, 64, 300), (64, 64, 101), (96, 81, 480)]


def _tol(shape, fast, general):
    from angularspectrum_heterogeneous_amd._lib import load  # noqa: F401  (the library must load)
    return general if shape in GEN else fast


def _params(shape, **over):
    from angularspectrum_heterogeneous_amd.workloads import params_kwargs
    kw = params_kwargs(dX=3.08e-4, dY=3.08e-4, dT=4e-8, propDist=0.02, dZmin=3.08e-4, boundaryFactor=0.2)
    kw.update(over)
    return O.OracleParams(**kw)


def _field(shape, seed=0, p0=2e5):
    from angularspectrum_heterogeneous_amd.workloads import axes, tone_burst
    nX, nY, nT = shape
    rng = np.random.default_rng(seed)
    x, y, t = axes(nX, 3.08e-4), axes(nY, 3.08e-4), axes(nT, 4e-8)
    ap = np.exp(-(x[:, None] ** 2 + y[None, :] ** 2) / (2 * (0.15 * nX * 3.08e-4) ** 2))
    f = ap[:, :, None] * tone_burst(t, 1e6, p0, 3)[None, None, :]
    f += 0.01 * p0 * rng.standard_normal(shape)
    return f.astype(np.float32)


def _plan(shape, p, dz):
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    plan = MarchPlan(shape, p)
    plan.set_dz(dz)
    return plan


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", SIZES[:2] + GEN[:2] + GEN[4:5])
@pytest.mark.parametrize("scheme", ["kt", "kt_minmod", "rusanov", "tvd"])
def test_nonlinear_step(shape, scheme):
    from angularspectrum_heterogeneous_amd import _lib
    p = _params(shape)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    u = _field(shape, p0=3e6)
    dz = 3.08e-4
    plan = _plan(shape, p, dz)
    code = {"kt": _lib.NL_KT, "kt_minmod": _lib.NL_KT_MINMOD, "rusanov": _lib.NL_RUSANOV,
            "tvd": _lib.NL_TVD}[scheme]
    got = plan.nonlinear_step(_dev(u), code, dz, beta_tvd=1.3).cpu().numpy()
    u64 = u.astype(np.float64)
    ref = {"kt": lambda: O.kt_flux(u64, N, dz, p.dT),
           "kt_minmod": lambda: O.kt_flux_minmod(u64, N, dz, p.dT),
           "rusanov": lambda: O.rusanov_flux_standard(u64, N, dz, p.dT),
           "tvd": lambda: O.rusanov_flux_tvd(u64, N, dz, p.dT, 1.3)}[scheme]()
    print(shape, scheme, rel_l2(got, ref), rel_l2(got - u, ref - u64))
    assert rel_l2(got, ref) < 1e-6
    assert rel_l2(got - u, ref - u64) < 1e-4  # the increment itself


@pytest.mark.parametrize("shape", SIZES + GEN)
def test_attenuation_step(shape):
    p = _params(shape)
    nX, nY, nT = shape
    dz = 3.08e-4
    u = _field(shape)
    plan = _plan(shape, p, dz)
    got = plan.attenuation_step(_dev(u), 0.5).cpu().numpy()
    _, afh, _, _, _ = O.precalculate_ad_pow2(p.alpha0, nX, nY, nT, p.dX, p.dY, dz, p.dT, p.c0, p.f0,
                                             p.attenPow, split_step=True, n_bins=nT // 2 + 1)
    ref = O.attenuation_step(u.astype(np.float64), afh)
    e = rel_l2(got, ref)
    print(shape, e)
    assert e < _tol(shape, 2e-6, 5e-6)


@pytest.mark.parametrize("shape", SIZES + GEN)
@pytest.mark.parametrize("half", [True, False])
def test_angular_spectrum_step(shape, half):
    p = _params(shape)
    nX, nY, nT = shape
    dz = 3.08e-4
    u = _field(shape)
    plan = _plan(shape, p, dz)
    got = plan.angular_spectrum_step(_dev(u), half).cpu().numpy()
    HH, HHh = O.precalculate_mas(nX, nY, nT, p.dX, p.dY, dz, p.dT, p.c0, split_step=True,
                                 adaptive_filtering=True, want_full=True)
    abl, ablh = O.precalculate_abl(nX, nY, nT, p.boundaryFactor, True, p.boundaryProfile)
    ref = O.angular_spectrum_step(u.astype(np.float64), HHh if half else HH, ablh if half else abl)
    e = rel_l2(got, ref)
    print(shape, half, e)
    assert e < _tol(shape, 2e-6, 5e-6)


@pytest.mark.parametrize("shape", SIZES[:3] + GEN)
def test_march_step_split_kt(shape):
    p = _params(shape)
    nX, nY, nT = shape
    dz = 3.08e-4
    u = _field(shape, p0=1e6)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    plan = _plan(shape, p, dz)
    got = plan.march_step(_dev(u)).cpu().numpy()
    ops = O.build_operators(nX, nY, nT, p, dz, np.float64)
    ref = O.march_step(u.astype(np.float64), ops, p, N, dz)
    e = rel_l2(got, ref)
    print(shape, e)
    assert e < _tol(shape, 5e-6, 1e-5)


@pytest.mark.parametrize("shape", SIZES[:3] + SIZES[4:] + GEN[:1] + GEN[3:4])
def test_march_step_nonlinearity_obliquity(shape):
    """The composite step with N scaled by the beam obliquity (ASS:956-965)."""
    p = _params(shape, useNonlinearityObliquity=True)
    nX, nY, nT = shape
    dz = 3.08e-4
    u = _field(shape, p0=1e6)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    plan = _plan(shape, p, dz)
    got = plan.march_step(_dev(u)).cpu().numpy()
    ops = O.build_operators(nX, nY, nT, p, dz, np.float64)
    obl = O.precalculate_obliquity_map(nX, nY, nT, p.dX, p.dY, p.dT, p.c0)
    ref = O.march_step(u.astype(np.float64), ops, p, N, dz, obl_map=obl)
    e = rel_l2(got, ref)
    print(shape, e)
    assert e < _tol(shape, 5e-6, 1e-5)


@pytest.mark.parametrize("shape", SIZES)
@pytest.mark.parametrize("launch", [(1, 0, 2048), (0, 1, 512), (1, 1, 4096)])
def test_march_step_launch_configs(shape, launch):
    """The autotuner's alternatives (radix orders of the x and y transforms, column grid)."""
    p = _params(shape)
    nX, nY, nT = shape
    dz = 3.08e-4
    u = _field(shape, p0=1e6)
    N = p.beta / (2 * p.c0 ** 3 * p.rho0)
    plan = _plan(shape, p, dz)
    plan.set_launch(*launch)
    got = plan.march_step(_dev(u)).cpu().numpy()
    ops = O.build_operators(nX, nY, nT, p, dz, np.float64)
    ref = O.march_step(u.astype(np.float64), ops, p, N, dz)
    e = rel_l2(got, ref)
    print(shape, launch, e)
    assert e < 5e-6


@pytest.mark.parametrize("shape", SIZES[:2] + SIZES[3:4] + GEN)
def test_freq_weighted_boundary(shape):
    p = _params(shape, useFreqWeightedBoundary=True)
    nX, nY, nT = shape
    u = _field(shape)
    plan = _plan(shape, p, 3.08e-4)
    got = plan.freq_weighted_boundary_step(_dev(u)).cpu().numpy()
    af = O.precalculate_abl_freq(nX, nY, nT, p.dT, p.c0, p.f0, p.boundaryFactor, p.boundaryProfile)
    ref = O.freq_weighted_boundary_step(u.astype(np.float64), af)
    e = rel_l2(got, ref)
    print(shape, e)
    assert e < _tol(shape, 2e-6, 5e-6)


@pytest.mark.parametrize("shape", SIZES[:2] + SIZES[3:4] + GEN)
def test_super_absorbing(shape):
    p = _params(shape, useSuperAbsorbing=True)
    nX, nY, nT = shape
    u = _field(shape)
    plan = _plan(shape, p, 3.08e-4)
    got = plan.super_absorbing_step(_dev(u)).cpu().numpy()
    wx = O.make_boundary_weights(nX, O.boundary_width(nX, p.boundaryFactor))
    wy = O.make_boundary_weights(nY, O.boundary_width(nY, p.boundaryFactor))
    ref = O.super_absorbing_step(u.astype(np.float64), wx.astype(np.float64), wy.astype(np.float64),
                                 p.dX, p.dY, p.dT, p.c0, p.superAbsorbingStrength)
    e = rel_l2(got, ref)
    print(shape, e)
    assert e < _tol(shape, 2e-6, 5e-6)


@pytest.mark.parametrize("shape", SIZES[:2] + GEN[:2] + GEN[3:4])
@pytest.mark.parametrize("with_amp", [False, True])
def test_phase_screen(shape, with_amp):
    p = _params(shape)
    nX, nY, nT = shape
    u = _field(shape)
    rng = np.random.default_rng(5)
    phase = rng.uniform(-3, 3, (nX, nY)).astype(np.float32)
    amp = rng.uniform(0.2, 1.2, (nX, nY)).astype(np.float32) if with_amp else None
    plan = _plan(shape, p, 3.08e-4)
    got = plan.apply_phase_screen(_dev(u), _dev(phase), _dev(amp) if with_amp else None).cpu().numpy()
    f0_bin = p.f0 / (1.0 / (nT * p.dT))
    ref = O.apply_phase_screen(u.astype(np.float64), phase.astype(np.float64), f0_bin,
                               None if amp is None else amp.astype(np.float64))
    e = rel_l2(got, ref)
    print(shape, with_amp, e)
    assert e < 5e-6
