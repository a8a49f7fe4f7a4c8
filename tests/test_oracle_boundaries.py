"""CPU pins of the oracle's boundary treatments and of the scheme dispatch.

* The super-absorbing step (ASS:185-272) on single-frequency plane waves, where
  the Engquist-Majda decomposition has a closed form: the central difference of
  exp(i k x) is i sin(k d)/d exp(i k x), so with coeff = -i c0/omega the
  "incoming" part at the left band is F (1 + s sin(k d)/(k d)) / 2 for a wave
  exp(i s k x).  A wave travelling towards a band's edge is removed (factor
  ~ 1 - strength w), one travelling away from it is kept; the y pass acts on the
  x-corrected spectrum.
* The published boundary/centre intensity ratios of validate_boundaries.py test 2
  (:92-168; README.md:11, :73): 0.0183 / 0.0135 / 0.0094 / 0.0075 for quadratic,
  Wendland, + frequency-weighted, + super-absorbing.
* The scheme dispatch (ASS:1454-1492) by a property of the schemes it selects: the
  standard Rusanov flux is twice the KT/TVD flux (-(uL^2+uR^2)/2 against -u^2/2),
  so a smooth pulse steepens twice as fast under it.
"""
import numpy as np
import pytest

from oracle import asm_oracle as O
from tests.conftest import rel_l2

C0, DT, NT, MBIN = 1540.0, 4e-8, 64, 3


def _plane_wave(n_axis, n_other, d, s, axis):
    """Real field cos(w t + s k pos) along `axis` (0 = x, 1 = y), uniform along the
    other axis; its rfft bin MBIN is (NT/2) exp(i s k pos) exactly.  s = +1 is
    p = g(t + pos/c0), a wave travelling towards -pos; s = -1 travels towards +pos."""
    w = 2 * np.pi * MBIN / (NT * DT)
    k = w / C0
    t = np.arange(NT) * DT
    pos = np.arange(n_axis) * d
    line = np.cos(w * t[None, :] + s * k * pos[:, None])          # (n_axis, NT)
    if axis == 0:
        f = np.broadcast_to(line[:, None, :], (n_axis, n_other, NT))
    else:
        f = np.broadcast_to(line[None, :, :], (n_other, n_axis, NT))
    F = (NT / 2) * np.exp(1j * s * k * pos)
    return np.ascontiguousarray(f), F, k


def _band_factor(F, k, d, s, w_lo, w_hi, strength):
    """Closed form of one SA pass along an axis for the plane wave F = A exp(i s k pos):
    the result is F times this factor (per index).  Interior: central difference
    i s sin(kd)/d F; ends: one-sided differences (exp(i s k d) - 1)/d F and
    (1 - exp(-i s k d))/d F."""
    n = len(F)
    dlog = np.full(n, 1j * s * np.sin(k * d) / d, dtype=complex)   # dF/F
    dlog[0] = (np.exp(1j * s * k * d) - 1) / d
    dlog[-1] = (1 - np.exp(-1j * s * k * d)) / d
    cd = (-1j / k) * dlog                                           # coeff dF / F, coeff = -i c0/omega
    return 1 - strength * w_lo * 0.5 * (1 + cd) - strength * w_hi * 0.5 * (1 - cd)


def _rfft_bin(g):
    G = np.fft.rfft(g, axis=2)
    rest = np.delete(G, MBIN, axis=2)
    return G[:, :, MBIN], float(np.abs(rest).max())


@pytest.mark.parametrize("kd", [0.05, 0.9])
@pytest.mark.parametrize("s", [1, -1])
def test_super_absorbing_x_pass_removes_the_outgoing_plane_wave(s, kd):
    nX, nY, bx, by, st = 24, 20, 6, 5, 0.8
    k = 2 * np.pi * MBIN / (NT * DT) / C0
    dX = dY = kd / k
    f, F, _ = _plane_wave(nX, nY, dX, s, 0)
    wx = O.make_boundary_weights(nX, bx).astype(np.float64)
    wy = O.make_boundary_weights(nY, by).astype(np.float64)
    g = O.super_absorbing_step(f, wx, wy, dX, dY, DT, C0, st)
    G, leak = _rfft_bin(g)
    assert leak < 1e-9 * NT                                     # one frequency in, one out
    ix = np.arange(nX)
    wl, wr = wx * (ix < nX // 2), wx * (ix >= nX // 2)
    fx = _band_factor(F, k, dX, s, wl, wr, st)
    # uniform along y: dF/dy = 0, so each y band removes strength * wy / 2 of the x-corrected F
    fy = 1 - st * wy * 0.5
    want = (F * fx)[:, None] * fy[None, :]
    np.testing.assert_allclose(G, want, rtol=1e-10, atol=1e-9 * NT)
    if kd < 0.1:
        # the physics: the band at the edge the wave travels towards takes out
        # ~strength*w of it, the band at the edge it comes from keeps it (to (k d)^2)
        towards = wl if s == 1 else wr
        inner = slice(1, nX - 1)
        np.testing.assert_allclose(fx[inner], (1 - st * towards)[inner], atol=st * kd ** 2)


@pytest.mark.parametrize("s", [1, -1])
def test_super_absorbing_y_pass_acts_on_the_x_corrected_spectrum(s):
    nX, nY, bx, by, st, kd = 20, 26, 5, 7, 0.6, 0.3
    k = 2 * np.pi * MBIN / (NT * DT) / C0
    dX = dY = kd / k
    f, F, _ = _plane_wave(nY, nX, dY, s, 1)
    wx = O.make_boundary_weights(nX, bx).astype(np.float64)
    wy = O.make_boundary_weights(nY, by).astype(np.float64)
    g = O.super_absorbing_step(f, wx, wy, dX, dY, DT, C0, st)
    G, leak = _rfft_bin(g)
    assert leak < 1e-9 * NT
    iy = np.arange(nY)
    fy = _band_factor(F, k, dY, s, wy * (iy < nY // 2), wy * (iy >= nY // 2), st)
    fx = 1 - st * wx * 0.5                      # uniform along x: the x pass scales by 1 - s w/2
    want = fx[:, None] * (F * fy)[None, :]      # y pass on the x-corrected spectrum: a product
    np.testing.assert_allclose(G, want, rtol=1e-10, atol=1e-9 * NT)
    # the corners tell the order apart: a y pass on the uncorrected spectrum would give
    # F (fx + fy - 1) there, which differs from F fx fy by F (1 - fx)(1 - fy)
    alt = F[None, :] * (fx[:, None] + fy[None, :] - 1)
    corner = (wx[:, None] > 0) & (wy[None, :] > 0)
    assert np.abs(alt - want)[corner].max() > 1e-2 * np.abs(F).max()


def _reflection_ic(nX, nY, nT, dX, dT, f0, p0, wX, wY, ncycles=1.5):
    """validate_boundaries.py:39-50: a square-aperture 1.5-cycle Gaussian burst."""
    x = np.arange(nX) * dX - (nX - 1) * dX / 2
    y = np.arange(nY) * dX - (nY - 1) * dX / 2
    t = np.arange(nT) * dT
    t -= np.mean(t)
    w0 = 2 * np.pi * f0
    ap = ((np.abs(x)[:, None] <= wX / 2) & (np.abs(y)[None, :] <= wY / 2)).astype(float)
    burst = np.exp(-(1.05 * t * w0 / (ncycles * np.pi)) ** 4) * np.sin(t * w0) * p0
    return ap[:, :, None] * burst[None, None, :]


REFLECTION_CASES = [
    ("quadratic", dict(boundaryProfile='quadratic'), 0.0183),
    ("wendland", dict(boundaryProfile='wendland'), 0.0135),
    ("freq_weighted", dict(boundaryProfile='wendland', useFreqWeightedBoundary=True), 0.0094),
    ("all_combined", dict(boundaryProfile='wendland', useFreqWeightedBoundary=True, useSuperAbsorbing=True,
                          superAbsorbingStrength=0.5), 0.0075),
]


def reflection_setup():
    """validate_boundaries.py:92-131 (test 2): 81 x 81 x 101, 3 MHz, lossless water, beta 0."""
    f0, c0 = 3e6, 1500.0
    dX = c0 / f0 / 5
    dT = dX / (5 * c0)
    nX = nY = 81
    field = _reflection_ic(nX, nY, 101, dX, dT, f0, 1e3, 8e-3, 8e-3)
    base = dict(dX=dX, dY=dX, dT=dT, c0=c0, rho0=1000, beta=0.0, alpha0=-1, f0=f0, propDist=3e-2,
                boundaryFactor=0.15, useSplitStep=True, useAdaptiveFiltering=False, useTVD=False, dZmin=dX * 4)
    return field, base


def reflection_ratio(pI_last):
    """validate_boundaries.py:134-141: boundary-band over centre intensity at the last z."""
    n = pI_last.shape[0]
    bw = int(n * 0.15)
    c = slice(bw, n - bw)
    ib = (pI_last[:bw, :].sum() + pI_last[-bw:, :].sum() + pI_last[c, :bw].sum() + pI_last[c, -bw:].sum())
    return float(ib / pI_last[c, c].sum())


def test_boundary_reflection_ratios_match_published():
    field, base = reflection_setup()
    got = []
    for label, kw, published in REFLECTION_CASES:
        res = O.angular_spectrum_solve(field, O.OracleParams(**base, **kw), dtype=np.float32)
        r = reflection_ratio(res[3][:, :, -1])
        print(f"{label}: {r:.4e} (published {published})")
        assert len(res[5]) == 75
        assert abs(r - published) < 0.0051 * published + 5e-5, (label, r, published)
        got.append(r)
    assert got[0] > got[1] > got[2] > got[3]


NT_NL, DT_NL = 512, 2e-8
N_WATER = 3.5 / (2 * 1540.0 ** 3 * 1000.0)


def _nl_burst():
    t = (np.arange(NT_NL) - (NT_NL - 1) / 2) * DT_NL
    u0 = 1e6 * np.exp(-(t / (80 * DT_NL)) ** 2) * np.sin(2 * np.pi * 1e6 * t)
    return np.broadcast_to(u0, (2, 2, NT_NL)).copy()


def _ops(split, hh=None, af=None):
    """Operators of the composite step; all-ones (identity linear parts) unless given."""
    s, sf = (2, 2, NT_NL), (2, 2, NT_NL // 2 + 1)
    hh = np.ones(s, complex) if hh is None else hh
    af = np.ones(sf, complex) if af is None else af
    if split:
        return {'HH_half': hh, 'abl_half': np.ones(s), 'afilt_half': af}
    return {'HH': hh, 'abl': np.ones(s), 'afilt': af}


def _params(**kw):
    return O.OracleParams(dX=3e-4, dY=3e-4, dT=DT_NL, c0=1540.0, f0=1e6, **kw)


def _march(p, z, dz, field=None):
    u = _nl_burst() if field is None else field
    ops = _ops(p.useSplitStep)
    for _ in range(int(round(z / dz))):
        u = O.march_step(u, ops, p, N_WATER, dz)
    return u[0, 0]


def _second_harmonic(u):
    f = np.fft.rfftfreq(NT_NL, DT_NL)
    return float(np.abs(np.fft.rfft(u))[np.argmin(np.abs(f - 2e6))])


def test_scheme_dispatch_by_flux_scale_and_order():
    """The composite step chosen by the dispatch (ASS:1454-1492), with identity linear
    operators so that only its nonlinear update acts, told apart by two properties:
    the flux scale (standard Rusanov's -(uL^2+uR^2)/2 is twice the -u^2/2 of KT and
    TVD, so it generates the second harmonic of a tone burst twice as fast) and the
    order in dZ (KT's SSP-RK2 is second order, the forward-Euler Rusanov forms first)."""
    cases = {'kt': _params(fluxScheme='kt'), 'tvd': _params(fluxScheme='rusanov', useTVD=True),
             'standard': _params(fluxScheme='rusanov', useTVD=False),
             'sequential': _params(fluxScheme='kt', useSplitStep=False)}
    z = 0.05
    h2, order = {}, {}
    for name, p in cases.items():
        ref = _march(p, z, z / 64)
        e4, e8 = (np.linalg.norm(_march(p, z, z / n) - ref) for n in (4, 8))
        order[name] = float(np.log2(e4 / e8))
        h2[name] = _second_harmonic(_march(p, z, z / 16))
    print(h2, order)
    assert 1.7 < h2['standard'] / h2['kt'] < 2.3
    assert 0.8 < h2['tvd'] / h2['kt'] < 1.2
    assert order['kt'] > 1.8
    assert order['tvd'] < 1.4 and order['standard'] < 1.4
    # the sequential step always takes standard Rusanov (ASS:874-880), whatever fluxScheme says
    np.testing.assert_allclose(_march(cases['sequential'], z, z / 16), _march(cases['standard'], z, z / 16),
                               rtol=0, atol=1e-6)


def test_sequential_step_order_propagator_then_nonlinear_then_attenuation():
    """march_step_sequential (ASS:874-880) applies the propagator, then the nonlinear
    update, then the attenuation filter.  With a low-pass attenuation filter the
    output holds no second harmonic (the filter acts after the update that makes it);
    with the same low-pass as the propagator it does (the update acts after it)."""
    p = _params(fluxScheme='rusanov', useSplitStep=False)
    f = np.fft.rfftfreq(NT_NL, DT_NL)
    lp_f = (f < 1.5e6).astype(complex)                                   # natural m order
    fs = np.fft.fftshift(np.fft.fftfreq(NT_NL, DT_NL))
    lp_t = (np.abs(fs) < 1.5e6).astype(complex)                          # fftshifted t order
    u = _nl_burst()
    hi = f > 1.7e6
    a_lp = O.march_step(u, _ops(False, af=np.broadcast_to(lp_f, (2, 2, len(f))).copy()), p, N_WATER, 2e-3)
    h_lp = O.march_step(u, _ops(False, hh=np.broadcast_to(lp_t, (2, 2, NT_NL)).copy()), p, N_WATER, 2e-3)
    spec_a = np.abs(np.fft.rfft(a_lp[0, 0]))
    spec_h = np.abs(np.fft.rfft(h_lp[0, 0]))
    base = np.abs(np.fft.rfft(u[0, 0])).max()
    assert spec_a[hi].max() < 1e-9 * base
    assert spec_h[hi].max() > 1e-3 * base
