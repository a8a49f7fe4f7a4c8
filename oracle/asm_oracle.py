"""CPU oracle: a NumPy/SciPy restatement of the reference's angular-spectrum march.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path may import this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

Every function restates one function of the reference module
``angular_spectrum_solver.py`` (cited below as ``ASS:<line>``) from its text.
Two precisions are supported:

* ``dtype=np.float64`` -- the parity oracle.  Operators stay complex128, the
  field stays float64 and all FFTs run in double precision.
* ``dtype=np.float32`` -- mirrors the reference's own dtype flow: operators are
  built in float64 and cast to complex64/float32 (ASS:1343-1351), the field is
  float32 (ASS:1408) and the FFTs run in single precision.  This mode is the
  timed CPU baseline ("reference algorithm restated on CPU").

The FFT library is ``scipy.fft`` (pocketfft) with ``workers`` threads.  The
reference's jitted kernels use ``jax.numpy.fft`` (XLA/ducc); both are exact DFTs
up to rounding, and both c2r transforms discard the imaginary part of the DC and
Nyquist bins, which the pipeline relies on (checked in the tests).

Each function is pinned in ``tests/test_oracle_pins.py`` and
``tests/test_oracle_boundaries.py`` by checks that do not retype its formula: closed
forms, invariants, the reference's known-answer test, published convergence orders
and the published boundary/centre intensity ratios (validate_boundaries.py test 2).
The super-absorbing step is pinned on plane waves (the outgoing wave is removed, the
y pass acts on the x-corrected spectrum); the scheme dispatch (``nonlinear_kind``,
ASS:1454-1492) by the flux scale and z order of the scheme it selects; the sequential
step's operator order (ASS:874-880) by low-pass operators.  Pinned only by
restatement: ``beta_tvd_for`` (ASS:1465-1466), whose value the TVD limiter cannot show
for the driver's beta in [1, 2] (min(beta, 1) = 1).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import scipy.fft as sfft

_WORKERS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def set_workers(n: int) -> None:
    """Thread count used by every scipy.fft call in this module."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def _cdt(dtype):
    return np.complex128 if np.dtype(dtype) == np.float64 else np.complex64


# ---------------------------------------------------------------------------
# Parameter container (ASS:35-76)
# ---------------------------------------------------------------------------
@dataclass
class OracleParams:
    """Same fields and defaults as the reference ``SolverParams`` (ASS:35-76)."""
    dX: float = 0.0
    dY: float = 0.0
    dT: float = 0.0
    c0: float = 1500.0
    rho0: float = 1000.0
    beta: float = 3.5
    alpha0: float = 0.5
    attenPow: float = 1.0
    f0: float = 5e6
    propDist: float = 0.08
    boundaryFactor: float = 0.2
    useSplitStep: bool = True
    useAdaptiveFiltering: bool = True
    useTVD: bool = True
    freqFilterThreshold: float = 0.05
    adaptiveFilterStrength: float = 0.7
    stabilityThreshold: float = 0.2
    stabilityRecoveryFactor: float = 0.15
    dZmin: float = 1e-3
    useBoundaryLayer: bool = True
    boundaryProfile: str = 'quadratic'
    useFreqWeightedBoundary: bool = False
    useSuperAbsorbing: bool = False
    superAbsorbingStrength: float = 0.8
    fluxScheme: str = 'rusanov'
    useObliquityCorrection: bool = True
    useNonlinearityObliquity: bool = False
    phaseScreens: object = None
    sourcePlanes: object = None


# ---------------------------------------------------------------------------
# Boundary tapers (ASS:82-138)
# ---------------------------------------------------------------------------
def taper_quadratic(N: int, n: int) -> np.ndarray:
    """ASS:82-91 ``ablvec``: 1 - s^2 on the n outer samples of each side."""
    s = np.zeros(N)
    i = np.arange(N)
    left = i < n
    right = i >= N - n
    s[left] = (n - i[left] - 1) / n
    s[right] = (i[right] - (N - n - 1)) / n
    out = np.zeros(N)
    out[left | right] = s[left | right] ** 2
    return 1.0 - out


def taper_wendland(N: int, n: int) -> np.ndarray:
    """ASS:94-110 ``ablvec_wendland``: 1 - 10 s^3 + 15 s^4 - 6 s^5, clipped to [0,1]."""
    vec = np.ones(N)
    i = np.arange(N)
    for mask, s in ((i < n, (n - i - 1) / n), (i >= N - n, (i - (N - n - 1)) / n)):
        ss = s[mask]
        vec[mask] = 1.0 - 10 * ss ** 3 + 15 * ss ** 4 - 6 * ss ** 5
    return np.clip(vec, 0.0, 1.0)


def taper(N: int, n: int, profile: str) -> np.ndarray:
    return taper_wendland(N, n) if profile == 'wendland' else taper_quadratic(N, n)


def boundary_width(n: int, factor: float) -> int:
    """ASS:121-123: ``max(round(n*factor), 1)`` with Python's half-to-even round."""
    return max(round(n * factor), 1)


def precalculate_abl(nX, nY, nT, boundary_factor=0.2, split_step=False,
                     profile='quadratic', dtype=np.float64):
    """ASS:116-131.  abl = ax(x) ay(y) at(t); abl_half = sqrt(abl) when split."""
    vx = taper(nX, boundary_width(nX, boundary_factor), profile)
    vy = taper(nY, boundary_width(nY, boundary_factor), profile)
    vt = taper(nT, boundary_width(nT, boundary_factor), profile)
    abl = (np.outer(vx, vy)[:, :, None] * vt[None, None, :])
    if np.dtype(dtype) == np.float32:
        abl = abl.astype(np.float32)          # ASS:127 casts before the sqrt
    abl_half = np.sqrt(abl) if split_step else None
    return abl, abl_half


def precalculate_identity_abl(nX, nY, nT, split_step=False, dtype=np.float64):
    """ASS:134-138."""
    abl = np.ones((nX, nY, nT), dtype=dtype)
    return abl, (abl.copy() if split_step else None)


def precalculate_abl_freq(nX, nY, nT, dT, c0, f0, boundary_factor=0.2,
                          profile='quadratic', dtype=np.float64):
    """ASS:144-179: abl_xy ** clip(f0 / max(f_m, f_1), 0.1, 10) on rfft bins."""
    vx = taper(nX, boundary_width(nX, boundary_factor), profile)
    vy = taper(nY, boundary_width(nY, boundary_factor), profile)
    abl_xy = np.outer(vx, vy)
    freqs = np.fft.rfftfreq(nT, dT)
    f_min = freqs[1] if len(freqs) > 1 else 1.0
    expo = np.clip(f0 / np.maximum(freqs, f_min), 0.1, 10.0)
    out = abl_xy[:, :, None] ** expo[None, None, :]
    # the reference stores this mask as float32 in both precisions (ASS:169, :1358)
    return out.astype(np.float32) if np.dtype(dtype) == np.float32 else out


def make_boundary_weights(N: int, n_bdy: int) -> np.ndarray:
    """ASS:275-284: linear ramp 1 at the edge -> 0 inside (float32 in the reference)."""
    w = np.zeros(N)
    i = np.arange(N)
    left = i < n_bdy
    right = i >= N - n_bdy
    w[left] = (n_bdy - i[left] - 1) / n_bdy
    w[right] = (i[right] - (N - n_bdy - 1)) / n_bdy
    return w.astype(np.float32)


# ---------------------------------------------------------------------------
# Super-absorbing boundary (ASS:185-272)
# ---------------------------------------------------------------------------
def super_absorbing_step(field, wx, wy, dX, dY, dT, c0, strength):
    """ASS:185-272: Engquist-Majda incoming-wave removal, x pass then y pass.

    ``field`` (nX,nY,nT) real.  The y pass works on the x-corrected spectrum.
    """
    rdt = field.dtype
    cdt = _cdt(rdt)
    nX, nY, nT = field.shape
    F = sfft.rfft(field, axis=2, workers=_WORKERS)
    nF = F.shape[2]
    freqs = np.fft.rfftfreq(nT, dT)
    omega = 2.0 * np.pi * freqs
    omega_min = 2.0 * np.pi * freqs[1] if nF > 1 else 1.0
    valid = np.abs(omega) > 0.5 * omega_min
    coeff = np.where(valid, -1j * c0 / np.where(valid, omega, 1.0), 0.0 + 0j).astype(cdt)[None, None, :]
    wx = np.asarray(wx).astype(rdt)
    wy = np.asarray(wy).astype(rdt)
    s = rdt.type(strength)

    dF = np.empty_like(F)
    dF[1:-1] = (F[2:] - F[:-2]) / rdt.type(2 * dX)
    dF[0] = (F[1] - F[0]) / rdt.type(dX)
    dF[-1] = (F[-1] - F[-2]) / rdt.type(dX)
    pin_l = rdt.type(0.5) * (F + coeff * dF)
    pin_r = rdt.type(0.5) * (F - coeff * dF)
    ix = np.arange(nX)
    wl = (wx * (ix < nX // 2))[:, None, None]
    wr = (wx * (ix >= nX // 2))[:, None, None]
    F = F - s * wl * pin_l
    F = F - s * wr * pin_r

    dF = np.empty_like(F)
    dF[:, 1:-1] = (F[:, 2:] - F[:, :-2]) / rdt.type(2 * dY)
    dF[:, 0] = (F[:, 1] - F[:, 0]) / rdt.type(dY)
    dF[:, -1] = (F[:, -1] - F[:, -2]) / rdt.type(dY)
    pin_b = rdt.type(0.5) * (F + coeff * dF)
    pin_t = rdt.type(0.5) * (F - coeff * dF)
    iy = np.arange(nY)
    wb = (wy * (iy < nY // 2))[None, :, None]
    wt = (wy * (iy >= nY // 2))[None, :, None]
    F = F - s * wb * pin_b
    F = F - s * wt * pin_t
    return sfft.irfft(F, n=nT, axis=2, workers=_WORKERS).astype(rdt)


# ---------------------------------------------------------------------------
# Phase screens (ASS:290-390)
# ---------------------------------------------------------------------------
def generate_phase_screen(nX, nY, dX, dY, c0, f0, correlation_length, speed_std,
                          thickness=None, seed=None):
    """ASS:290-349: Gaussian random sound-speed layer -> phase map (float32)."""
    if thickness is None:
        thickness = correlation_length
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nX, dX) * 2 * np.pi
    ky = np.fft.fftfreq(nY, dY) * 2 * np.pi
    K2 = kx[:, None] ** 2 + ky[None, :] ** 2
    spectrum = np.exp(-K2 * correlation_length ** 2 / 4.0)
    noise = rng.standard_normal((nX, nY)) + 1j * rng.standard_normal((nX, nY))
    delta_c = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spectrum)))
    delta_c = delta_c / (np.std(delta_c) + 1e-30) * speed_std
    c_map = c0 + delta_c
    phase = 2 * np.pi * f0 * thickness * (1.0 / c_map - 1.0 / c0)
    return phase.astype(np.float32), c_map.astype(np.float32)


def apply_phase_screen(field, phase_screen, f0_bin, amplitude_screen=None):
    """ASS:352-390: F *= exp(-i phi m/f0b) * clip(A,1e-6,1)^(m/f0b) on rfft bins."""
    rdt = field.dtype
    nT = field.shape[2]
    F = sfft.rfft(field, axis=2, workers=_WORKERS)
    nF = F.shape[2]
    fs = (np.arange(nF, dtype=rdt) / rdt.type(max(f0_bin, 1.0)))[None, None, :]
    ps = np.asarray(phase_screen).astype(rdt)[:, :, None]
    F = F * np.exp(-1j * (ps * fs)).astype(F.dtype)
    if amplitude_screen is not None:
        amp = np.clip(np.asarray(amplitude_screen).astype(rdt), 1e-6, 1.0)[:, :, None]
        F = F * np.power(amp, fs).astype(rdt)
    return sfft.irfft(F, n=nT, axis=2, workers=_WORKERS).astype(rdt)


# ---------------------------------------------------------------------------
# Wavenumber grids (ASS:403-408, 488-491)
# ---------------------------------------------------------------------------
def centered_grid(n: int, d: float, c0: float = None) -> np.ndarray:
    """(n-1)-normalised, mean-subtracted wavenumber grid (ASS:403-408); with ``c0``
    the ASM t grid, divided by c0 before the mean is taken, as ASS:403-404 does.

    Index j of the result multiplies the fftshifted bin j, i.e. natural FFT
    bin ``(j - n//2) % n``.
    """
    k = np.linspace(0, n - 1, n) / (n - 1) / d * 2 * np.pi
    if c0 is not None:
        k = k / c0
    return k - np.mean(k)


def kfilter_mask(kk, dX, dY, dZ, dT, c0, kt, adaptive, threshold, strength):
    """ASS:411-433: cosine roll-off of the transverse wavenumber."""
    kmax = min(np.pi / dX, np.pi / dY)
    nz = kt[kt != 0]
    lambda_char = c0 / np.mean(np.abs(nz)) if nz.size else c0 / (1.0 / dT)
    norm_step = dZ / lambda_char
    if adaptive:
        cutoff = threshold * (1 + strength * norm_step * 5)
        k_max = min(kmax * (1 - cutoff), kmax * 0.95)
        k_start = k_max * 0.8
    else:
        k_max = kmax * 0.95
        k_start = k_max * 0.9
    k_tr = np.sqrt(kk)
    filt = np.ones_like(k_tr)
    trans = (k_tr > k_start) & (k_tr <= k_max)
    if np.any(trans):
        pos = (k_tr[trans] - k_start) / (k_max - k_start)
        filt[trans] = 0.5 * (1 + np.cos(np.pi * pos))
    filt[k_tr > k_max] = 0.0
    return filt, (k_start, k_max)


def precalculate_mas(nX, nY, nT, dX, dY, dZ, dT, c0, split_step=False,
                     adaptive_filtering=False, filter_threshold=0.05,
                     filter_strength=0.7, want_full=True):
    """ASS:396-462: analytic-signal angular-spectrum propagator (complex128).

    Returns (HH, HH_half) in fftshifted (kx, ky, kt) order, with fftshifted
    t-indices 0..nT//2 zeroed and the rest doubled.  ``want_full=False`` skips
    HH when only the split-step half operator is needed (the reference builds
    both; the unused one does not change any result).
    """
    kt = centered_grid(nT, dT, c0)
    kx = centered_grid(nX, dX)
    ky = centered_grid(nY, dY)
    kk = kx[:, None] ** 2 + ky[None, :] ** 2
    filt, _ = kfilter_mask(kk, dX, dY, dZ, dT, c0, kt, adaptive_filtering,
                           filter_threshold, filter_strength)
    HH = np.zeros((nX, nY, nT), dtype=np.complex128) if (want_full or not split_step) else None
    HHh = np.zeros((nX, nY, nT), dtype=np.complex128) if split_step else None
    for m in range(nT // 2 + 1, nT):          # bins below are zeroed anyway (ASS:455-459)
        k = kt[m]
        for arr, dz in ((HH, dZ), (HHh, dZ / 2)):
            if arr is None:
                continue
            H2 = np.exp(dz * (1j * k - np.sqrt(kk - k ** 2 + 0j)))
            H1 = np.exp(dz * (1j * k - 1j * np.sqrt(k ** 2 - kk + 0j)))
            arr[:, :, m] = np.where(kk < k ** 2, H1, H2) * filt
    for arr in (HH, HHh):
        if arr is not None:
            arr *= 2
    return HH, HHh


# ---------------------------------------------------------------------------
# Attenuation / dispersion (ASS:468-592)
# ---------------------------------------------------------------------------
def build_oblique_atten_filter(alpha, alphaStar, nX, nY, nT, dX, dY, dZ, c0, dT,
                               scale=1.0, obliquity=True, clip_negative=True,
                               n_bins=None):
    """ASS:468-520, restricted to the first ``n_bins`` temporal bins.

    The reference builds all nT bins and the driver slices the rfft support
    (ASS:1331-1335); the remaining bins are never used.
    """
    n_bins = nT if n_bins is None else n_bins
    kx = centered_grid(nX, dX)
    ky = centered_grid(nY, dY)
    kk = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.arange(nT) / (nT * dT)
    k = 2 * np.pi * f / c0
    out = np.ones((nX, nY, n_bins), dtype=np.complex128)
    eps = 1e-6
    for m in range(n_bins):
        km = k[m]
        if km == 0:
            continue
        base = -(alpha[m] + 1j * alphaStar[m]) * dZ * scale
        if obliquity:
            kz_sq = km ** 2 - kk
            prop = kz_sq > 0
            kz_safe = np.where(prop, np.maximum(np.sqrt(np.maximum(kz_sq, 0.0)), np.abs(km) * eps), 1.0)
            obl = np.where(prop, np.abs(km) / kz_safe, 0.0)
            fac = np.where(prop, np.exp(base * obl), 1.0)
        else:
            fac = np.broadcast_to(np.exp(base), kk.shape)
        out[:, :, m] = fac
    if clip_negative:
        out = np.where(np.real(out) < 0, 0, out)
    return out


def water_alpha(alpha0, nT, dT, f0):
    """ASS:523-538: f^2 attenuation with the tan(pi) KK term."""
    f = np.arange(nT) / (nT * dT)
    conv = alpha0 / 1e12 * 1e2 / (20 * np.log10(np.e))
    alpha = conv * f ** 2
    a0 = (conv / (2 * np.pi) ** 2) * np.tan(np.pi) * ((2 * np.pi * f) ** 1 - (2 * np.pi * f0) ** 1)
    astar = 2 * np.pi * a0 * f
    astar[0] = 0
    return alpha, astar


def powerlaw_alpha(alpha0, nT, dT, f0, pw):
    """ASS:555-575: power-law attenuation with Kramers-Kronig dispersion."""
    f = np.arange(nT) / (nT * dT)
    f_safe = f.copy()
    if nT > 1:
        f_safe[0] = f_safe[1] / 2
    conv = alpha0 / (1e6 ** pw) * 1e2 / (20 * np.log10(np.e))
    alpha = conv * f ** pw
    if pw % 2 == 1:
        a0 = (-2 * conv / ((2 * np.pi) ** pw) / np.pi) * (np.log(2 * np.pi * f_safe) - np.log(2 * np.pi * f0))
    else:
        a0 = (conv / (2 * np.pi) ** pw) * np.tan(np.pi * pw / 2) * \
             ((2 * np.pi * f) ** (pw - 1) - (2 * np.pi * f0) ** (pw - 1))
    astar = 2 * np.pi * a0 * f
    astar[0] = 0
    return alpha, astar, f, f_safe


def precalculate_ad(alpha0, nX, nY, nT, dX, dY, dZ, dT, c0, f0, split_step=False,
                    obliquity=True, n_bins=None):
    """ASS:523-549 (water branch, negative-real clipping on)."""
    alpha, astar = water_alpha(alpha0, nT, dT, f0)
    full = build_oblique_atten_filter(alpha, astar, nX, nY, nT, dX, dY, dZ, c0, dT,
                                      1.0, obliquity, True, n_bins)
    half = build_oblique_atten_filter(alpha, astar, nX, nY, nT, dX, dY, dZ, c0, dT,
                                      0.5, obliquity, True, n_bins) if split_step else None
    return full, half


def precalculate_ad_pow2(alpha0, nX, nY, nT, dX, dY, dZ, dT, c0, f0, pw, split_step=False,
                         obliquity=True, n_bins=None):
    """ASS:555-592 (no clipping).  Returns (afilt, afilt_half, f, alpha, dispersion)."""
    alpha, astar, f, f_safe = powerlaw_alpha(alpha0, nT, dT, f0, pw)
    full = build_oblique_atten_filter(alpha, astar, nX, nY, nT, dX, dY, dZ, c0, dT,
                                      1.0, obliquity, False, n_bins)
    half = build_oblique_atten_filter(alpha, astar, nX, nY, nT, dX, dY, dZ, c0, dT,
                                      0.5, obliquity, False, n_bins) if split_step else None
    dispersion = 1.0 / ((1.0 / c0) + (astar / (2 * np.pi * f_safe)))
    return full, half, f, alpha, dispersion


# ---------------------------------------------------------------------------
# Beam-averaged nonlinear obliquity (ASS:598-643)
# ---------------------------------------------------------------------------
def precalculate_obliquity_map(nX, nY, nT, dX, dY, dT, c0):
    """ASS:598-627: k/k_z per mode on rfft bins (float32), 0 evanescent and at DC."""
    kx = centered_grid(nX, dX)
    ky = centered_grid(nY, dY)
    kk = kx[:, None] ** 2 + ky[None, :] ** 2
    nF = nT // 2 + 1
    k_tf = 2 * np.pi * (np.arange(nF) / (nT * dT)) / c0
    out = np.zeros((nX, nY, nF), dtype=np.float32)
    for m in range(nF):
        km = k_tf[m]
        if km == 0:
            continue
        kz_sq = km ** 2 - kk
        prop = kz_sq > 0
        kz_safe = np.where(prop, np.maximum(np.sqrt(np.maximum(kz_sq, 0.0)), np.abs(km) * 1e-6), 1.0)
        out[:, :, m] = np.where(prop, np.abs(km) / kz_safe, 0.0).astype(np.float32)
    return out


def beam_obliquity_scalar(field, obl_map):
    """ASS:630-643: power-weighted <k/k_z> of the field's plane-wave spectrum."""
    F = sfft.rfft(field, axis=2, workers=_WORKERS)
    F = np.fft.fftshift(sfft.fft2(F, axes=(0, 1), workers=_WORKERS), axes=(0, 1))
    P = (F.real ** 2 + F.imag ** 2).astype(obl_map.dtype if field.dtype == np.float32 else np.float64)
    om = obl_map.astype(P.dtype)
    num = np.sum(P * om)
    den = np.sum(P * (om > 0).astype(P.dtype)) + 1e-30
    return num / den


# ---------------------------------------------------------------------------
# Linear step kernels (ASS:649-685)
# ---------------------------------------------------------------------------
def angular_spectrum_step(field, HH, abl):
    """ASS:649-655: Re(ifftn(ifftshift(fftshift(fftn(p)) HH))) * abl."""
    F = sfft.fftn(field, workers=_WORKERS)
    F = np.fft.ifftshift(np.fft.fftshift(F) * HH)
    out = np.real(sfft.ifftn(F, workers=_WORKERS))
    return (out * abl).astype(field.dtype)


def freq_weighted_boundary_step(field, abl_freq):
    """ASS:658-664."""
    nT = field.shape[2]
    F = sfft.rfft(field, axis=2, workers=_WORKERS) * abl_freq
    return sfft.irfft(F, n=nT, axis=2, workers=_WORKERS).astype(field.dtype)


def attenuation_step(field, afilt):
    """ASS:667-685: rfft_t, fft2_xy, centred multiply, back.  ``afilt`` is
    (nX, nY, nT//2+1) in fftshifted (kx, ky) order."""
    nT = field.shape[2]
    F = sfft.rfft(field, axis=2, workers=_WORKERS)
    F = np.fft.fftshift(sfft.fft2(F, axes=(0, 1), workers=_WORKERS), axes=(0, 1)) * afilt
    F = sfft.ifft2(np.fft.ifftshift(F, axes=(0, 1)), axes=(0, 1), workers=_WORKERS)
    return sfft.irfft(F, n=nT, axis=2, workers=_WORKERS).astype(field.dtype)


# ---------------------------------------------------------------------------
# Nonlinear (Burgers) steps along t (ASS:688-868)
# ---------------------------------------------------------------------------
def _s(x, ref):
    """Python scalar -> the field's dtype (JAX weak typing keeps float32 ops float32)."""
    return ref.dtype.type(x)


def rusanov_flux_standard(field, N, dZ, dT):
    """ASS:688-696: forward-Euler Rusanov with flux -(uL^2+uR^2)/2 - a (uR-uL)."""
    u = field
    lam = np.maximum(np.abs(u[..., :-1]), np.abs(u[..., 1:]))
    fh = -(u[..., :-1] ** 2 + u[..., 1:] ** 2) / _s(2, u) - lam * (u[..., 1:] - u[..., :-1])
    out = u.copy()
    out[..., 1:-1] += -_s(N * dZ / dT, u) * (fh[..., 1:] - fh[..., :-1])
    return out


def rusanov_flux_tvd(field, N, dZ, dT, beta_tvd):
    """ASS:699-724: Rusanov with the generalised-minmod TVD limiter."""
    u = field
    uL, uR = u[..., :-1], u[..., 1:]
    a = np.maximum(np.abs(uL), np.abs(uR))
    half = _s(0.5, u)
    fh = half * (-half * (uL ** 2 + uR ** 2)) - half * a * (uR - uL)
    fd = fh[..., 1:] - fh[..., :-1]
    dm = u[..., 1:-1] - u[..., :-2]
    dp = u[..., 2:] - u[..., 1:-1]
    eps = _s(1e-30, u)
    b = _s(beta_tvd, u)
    one = _s(1.0, u)
    zero = _s(0.0, u)
    with np.errstate(over='ignore', divide='ignore', invalid='ignore'):
        r = dp / (dm + eps * np.sign(dm + eps))
        ri = dm / (dp + eps * np.sign(dp + eps))
    pp = np.maximum(zero, np.minimum(np.minimum(b * r, one), np.minimum(r, b)))
    pm = np.maximum(zero, np.minimum(np.minimum(b * ri, one), np.minimum(ri, b)))
    out = u.copy()
    out[..., 1:-1] += -_s(N * dZ / dT, u) * (half * (pp + pm) * fd)
    return out


def minmod(a, b):
    """ASS:727-731."""
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0).astype(a.dtype)


def minmod3(a, b, c):
    """ASS:734-741."""
    same = (a * b > 0) & (b * c > 0)
    m = np.minimum(np.minimum(np.abs(a), np.abs(b)), np.abs(c))
    return np.where(same, np.sign(a) * m, 0.0).astype(a.dtype)


def mc_limiter(dm, dp):
    """ASS:744-755: minmod(2 dm, (dm+dp)/2, 2 dp)."""
    two = _s(2, dm)
    return minmod3(two * dm, _s(0.5, dm) * (dm + dp), two * dp)


def kt_rhs(field, N, dT, limiter='mc'):
    """ASS:758-797 (MC) and ASS:800-830 (minmod): semi-discrete KT rhs."""
    u = field
    dm = u[..., 1:-1] - u[..., :-2]
    dp = u[..., 2:] - u[..., 1:-1]
    sm = mc_limiter(dm, dp) if limiter == 'mc' else minmod(dm, dp)
    sigma = np.zeros_like(u)
    sigma[..., 1:-1] = sm
    half = _s(0.5, u)
    um = u[..., :-1] + half * sigma[..., :-1]
    up = u[..., 1:] - half * sigma[..., 1:]
    zero = _s(0.0, u)
    ap = np.maximum(np.maximum(-um, -up), zero)
    am = np.minimum(np.minimum(-um, -up), zero)
    fm = -half * um ** 2
    fp = -half * up ** 2
    den = ap - am
    small = np.abs(den) < 1e-14
    sden = np.where(small, _s(1.0, u), den)
    kt = (ap * fm - am * fp - ap * am * (up - um)) / sden
    lf = half * (fm + fp) - half * np.maximum(np.abs(um), np.abs(up)) * (up - um)
    fh = np.where(small, lf, kt)
    rhs = np.zeros_like(u)
    rhs[..., 1:-1] = -_s(N / dT, u) * (fh[..., 1:] - fh[..., :-1])
    return rhs


def kt_flux(field, N, dZ, dT, limiter='mc'):
    """ASS:833-839: SSP-RK2 (Heun) in z over the KT rhs."""
    dz = _s(dZ, field)
    half = _s(0.5, field)
    r0 = kt_rhs(field, N, dT, limiter)
    s1 = field + dz * r0
    r1 = kt_rhs(s1, N, dT, limiter)
    return half * field + half * (s1 + dz * r1)


def kt_flux_minmod(field, N, dZ, dT):
    """ASS:862-868."""
    return kt_flux(field, N, dZ, dT, limiter='minmod')


def kt_flux_adaptive(field, N, dZ, dT, cfl_target=0.5):
    """ASS:842-859: global-max CFL sub-cycling of ``kt_flux``."""
    u_max = float(np.max(np.abs(field)))
    cfl = N * u_max * dZ / dT
    if cfl <= cfl_target or u_max < 1e-30:
        return kt_flux(field, N, dZ, dT)
    n_sub = int(math.ceil(cfl / cfl_target))
    for _ in range(n_sub):
        field = kt_flux(field, N, dZ / n_sub, dT)
    return field


def kt_flux_subcycled_columns(field, N, dZ, dT, cfl_target=0.5):
    """Per-column variant of ``kt_flux_adaptive``: every (x, y) column picks its
    own n_sub = ceil(N max|u_col| dZ / dT / cfl_target) (1 when at or below the
    target).  Equal to ``kt_flux`` wherever the CFL number is <= the target."""
    out = np.empty_like(field)
    flat_in = field.reshape(-1, field.shape[-1])
    flat_out = out.reshape(-1, field.shape[-1])
    for c in range(flat_in.shape[0]):
        col = flat_in[c:c + 1]
        umax = float(np.max(np.abs(col)))
        cfl = N * umax * dZ / dT
        n_sub = 1 if (cfl <= cfl_target or umax < 1e-30) else int(math.ceil(cfl / cfl_target))
        for _ in range(n_sub):
            col = kt_flux(col, N, dZ / n_sub, dT)
        flat_out[c] = col[0]
    return out


# ---------------------------------------------------------------------------
# Composite march steps (ASS:874-965)
# ---------------------------------------------------------------------------
def march_step_sequential(field, HH, abl, afilt, N, dZ, dT, obl_map=None):
    """ASS:874-880; with ``obl_map`` the obliquity variant ASS:923-929."""
    field = angular_spectrum_step(field, HH, abl)
    Ne = N * beam_obliquity_scalar(field, obl_map) if obl_map is not None else N
    field = rusanov_flux_standard(field, Ne, dZ, dT)
    return attenuation_step(field, afilt)


def march_step_split(field, HHh, ablh, afilth, N, dZ, dT, nonlinear, obl_map=None):
    """ASS:883-913 (the three Strang variants differ only in ``nonlinear(u, N)``);
    with ``obl_map`` the obliquity variants ASS:932-965, whose nonlinear step
    takes N scaled by the beam's <k/k_z> after the first half step."""
    field = attenuation_step(field, afilth)
    field = angular_spectrum_step(field, HHh, ablh)
    Ne = N * beam_obliquity_scalar(field, obl_map) if obl_map is not None else N
    field = nonlinear(field, Ne)
    field = angular_spectrum_step(field, HHh, ablh)
    return attenuation_step(field, afilth)


def march_step_split_kt(field, HHh, ablh, afilth, N, dZ, dT, obl_map=None):
    """ASS:905-913 (ASS:956-965 with ``obl_map``)."""
    return march_step_split(field, HHh, ablh, afilth, N, dZ, dT,
                            lambda u, n: kt_flux(u, n, dZ, dT), obl_map)


def march_step_split_tvd(field, HHh, ablh, afilth, N, dZ, dT, beta_tvd, obl_map=None):
    """ASS:894-902 (ASS:944-953 with ``obl_map``)."""
    return march_step_split(field, HHh, ablh, afilth, N, dZ, dT,
                            lambda u, n: rusanov_flux_tvd(u, n, dZ, dT, beta_tvd), obl_map)


def march_step_split_standard(field, HHh, ablh, afilth, N, dZ, dT, obl_map=None):
    """ASS:883-891 (ASS:932-941 with ``obl_map``)."""
    return march_step_split(field, HHh, ablh, afilth, N, dZ, dT,
                            lambda u, n: rusanov_flux_standard(u, n, dZ, dT), obl_map)


# ---------------------------------------------------------------------------
# Bowl source planes (ASS:1080-1211)
# ---------------------------------------------------------------------------
def bowl_plane_geometry(xaxis, yaxis, c0, radius, roc, dZ, inner_radius=0.0,
                        focus=None, element_delays=None, n_elements=1):
    """ASS:1126-1166: aperture, bowl depth map, focusing delays and slice edges.

    Returns (aperture (bool), bowl_z, delays_full, z_edges, bowl_depth).
    """
    if focus is None:
        focus = roc
    R = roc
    XX, YY = np.meshgrid(np.asarray(xaxis, dtype=np.float64), np.asarray(yaxis, dtype=np.float64), indexing='ij')
    r2 = XX ** 2 + YY ** 2
    r = np.sqrt(r2)
    ap = (r2 <= radius ** 2) & (r2 >= inner_radius ** 2)
    bowl_z = R - np.sqrt(R ** 2 - np.minimum(r2, R ** 2 - 1e-20))
    bowl_z[~ap] = 0.0
    depth = float(np.max(bowl_z[ap]))
    dist = np.sqrt(r2 + (focus - bowl_z) ** 2)
    delays = np.zeros_like(r)
    delays[ap] = dist[ap] / c0
    if n_elements > 1:
        dr = (radius - inner_radius) / n_elements
        for ie in range(n_elements):
            m = (r >= inner_radius + ie * dr) & (r < inner_radius + (ie + 1) * dr) & ap
            if m.any():
                delays[m] = delays[m].mean()
                if element_delays is not None:
                    delays[m] += np.asarray(element_delays)[ie]
    delays[ap] = delays[ap].max() - delays[ap]
    n_slices = max(1, int(np.ceil(depth / dZ)))
    z_edges = np.linspace(0, depth, n_slices + 1)
    return ap, bowl_z, delays, z_edges, depth


def tone_burst(t_minus_d, f0, p0, ncycles, dur):
    """Gaussian-enveloped sinusoid used by every source builder (ASS:1195-1200)."""
    w0 = 2 * np.pi * f0
    env = np.exp(-(1.05 * t_minus_d * w0 / (ncycles * np.pi)) ** (2 * dur))
    return env * np.sin(w0 * t_minus_d) * p0


def make_bowl_source_planes(xaxis, yaxis, taxis, f0, c0, p0, radius, roc, dZ,
                            ncycles=3, dur=2, inner_radius=0.0, focus=None,
                            element_delays=None, n_elements=1):
    """ASS:1080-1211: dense (z_mid, float32 (nX,nY,nT)) slices, and bowl depth."""
    ap, bowl_z, dfull, z_edges, depth = bowl_plane_geometry(
        xaxis, yaxis, c0, radius, roc, dZ, inner_radius, focus, element_delays, n_elements)
    taxis = np.asarray(taxis, dtype=np.float64)
    planes = []
    n_slices = len(z_edges) - 1
    for i in range(n_slices):
        lo, hi = z_edges[i], z_edges[i + 1]
        m = ap & (bowl_z >= lo) & ((bowl_z <= hi) if i == n_slices - 1 else (bowl_z < hi))
        if not m.any():
            continue
        d = dfull.copy()
        d[m] -= bowl_z[m] / c0
        d[~m] = 0.0
        fld = m.astype(np.float32)[:, :, None] * tone_burst(taxis[None, None, :] - d[:, :, None],
                                                           f0, p0, ncycles, dur)
        planes.append((0.5 * (lo + hi), fld.astype(np.float32)))
    return planes, depth


def make_bowl_source(xaxis, yaxis, taxis, f0, c0, p0, radius, roc, ncycles=3, dur=2,
                     inner_radius=0.0, focus=None, element_delays=None, n_elements=1):
    """ASS:971-1077: one-plane bowl source, dense float32 (nX, nY, nT)."""
    if focus is None:
        focus = roc
    XX, YY = np.meshgrid(np.asarray(xaxis, dtype=np.float64), np.asarray(yaxis, dtype=np.float64), indexing='ij')
    r2 = XX ** 2 + YY ** 2
    r = np.sqrt(r2)
    ap = ((r2 <= radius ** 2) & (r2 >= inner_radius ** 2)).astype(np.float32)
    zb = roc - np.sqrt(roc ** 2 - np.minimum(r2, roc ** 2 - 1e-20))
    dist = np.sqrt(r2 + (focus - zb) ** 2)
    on = ap > 0
    delays = np.zeros_like(r)
    delays[on] = dist[on] / c0
    if n_elements > 1:
        dr = (radius - inner_radius) / n_elements
        for ie in range(n_elements):
            m = (r >= inner_radius + ie * dr) & (r < inner_radius + (ie + 1) * dr) & on
            if m.any():
                delays[m] = delays[m].mean()
                if element_delays is not None:
                    delays[m] += np.asarray(element_delays)[ie]
    delays[on] = delays[on].max() - delays[on]
    taxis = np.asarray(taxis, dtype=np.float64)
    fld = ap[:, :, None] * tone_burst(taxis[None, None, :] - delays[:, :, None], f0, p0, ncycles, dur)
    return fld.astype(np.float32)


# ---------------------------------------------------------------------------
# Transcranial inputs (validate_transcranial.py)
# ---------------------------------------------------------------------------
def skull_to_phase_screens(c_map, dz_screen, c0, f0, rho_map=None, rho0=1000.0, alpha_bone=8.0,
                           alpha_tissue=0.5):
    """validate_transcranial.py:166-238: [(z, phase, amplitude)] of the non-trivial slices."""
    out = []
    for iz in range(c_map.shape[2]):
        c = c_map[:, :, iz]
        phase = 2 * np.pi * f0 * dz_screen * (1.0 / c - 1.0 / c0)
        bone = np.clip((c - 1540.0) / (2900.0 - 1540.0), 0, 1)
        alpha = alpha_tissue + bone * (alpha_bone - alpha_tissue)
        amp = np.exp(-(alpha * (f0 / 1e6) * 0.1151 / 0.01) * dz_screen)
        if rho_map is not None:
            zl = rho_map[:, :, iz] * c
            amp = amp * np.abs(2.0 * zl / (zl + rho0 * c0))
        if np.max(np.abs(phase)) > 1e-6 or np.min(amp) < 0.999:
            out.append((iz * dz_screen, phase.astype(np.float32), amp.astype(np.float32)))
    return out


def build_initial_condition(positions, element_ids, xaxis, yaxis, t, f0, c0, p0, focal_depth, ncycles=2):
    """validate_transcranial.py:261-317, with its per-pixel element loop."""
    nX, nY = len(xaxis), len(yaxis)
    dX = xaxis[1] - xaxis[0]
    iy_raw = np.round(positions[:, 0] / dX).astype(int)
    iz_raw = np.round(positions[:, 1] / dX).astype(int)
    iyc = iy_raw - int(np.mean(iy_raw)) + nX // 2
    izc = iz_raw - int(np.mean(iz_raw)) + nY // 2
    dist = np.sqrt(focal_depth ** 2 + (iyc * dX - (nX // 2) * dX) ** 2 + (izc * dX - (nY // 2) * dX) ** 2)
    delays = (dist.max() - dist) / c0
    half = max(1, round(750e-6 / dX)) // 2
    ap = np.zeros((nX, nY), dtype=np.float32)
    tt = np.zeros((nX, nY), dtype=np.float32)
    for m in range(len(element_ids)):
        for di in range(-half, half + 1):
            for dj in range(-half, half + 1):
                yi, zi = iyc[m] + di, izc[m] + dj
                if 0 <= yi < nX and 0 <= zi < nY:
                    ap[yi, zi] = 1.0
                    tt[yi, zi] = delays[m]
    fld = tone_burst(np.asarray(t, np.float64)[None, None, :] - tt[:, :, None], f0, p0, ncycles, 2)
    return (ap[:, :, None] * fld).astype(np.float32)


# ---------------------------------------------------------------------------
# Driver (ASS:1217-1567)
# ---------------------------------------------------------------------------
class StabilityError(RuntimeError):
    """Raised where the reference would restart forever with an unchanged dZ."""


def build_operators(nX, nY, nT, p: OracleParams, dZ, dtype):
    """ASS:1301-1338 + casts of ASS:1343-1351 (only what the march uses)."""
    split = p.useSplitStep
    cdt = _cdt(dtype)
    HH, HHh = precalculate_mas(nX, nY, nT, p.dX, p.dY, dZ, p.dT, p.c0, split_step=split,
                               adaptive_filtering=p.useAdaptiveFiltering,
                               filter_threshold=p.freqFilterThreshold,
                               filter_strength=p.adaptiveFilterStrength, want_full=not split)
    if p.useBoundaryLayer:
        abl, ablh = precalculate_abl(nX, nY, nT, p.boundaryFactor, split_step=split,
                                     profile=p.boundaryProfile, dtype=dtype)
    else:
        abl, ablh = precalculate_identity_abl(nX, nY, nT, split_step=split, dtype=dtype)
    nF = nT // 2 + 1
    if p.alpha0 < 0:
        af, afh = precalculate_ad(2.17e-3, nX, nY, nT, p.dX, p.dY, dZ, p.dT, p.c0, p.f0,
                                  split_step=split, obliquity=p.useObliquityCorrection, n_bins=nF)
    else:
        af, afh, _, _, _ = precalculate_ad_pow2(p.alpha0, nX, nY, nT, p.dX, p.dY, dZ, p.dT, p.c0,
                                                p.f0, p.attenPow, split_step=split,
                                                obliquity=p.useObliquityCorrection, n_bins=nF)
    ops = {}
    if split:
        ops['HH_half'] = HHh.astype(cdt)
        ops['abl_half'] = ablh.astype(dtype)
        ops['afilt_half'] = afh.astype(cdt)
    else:
        ops['HH'] = HH.astype(cdt)
        ops['abl'] = abl.astype(dtype)
        ops['afilt'] = af.astype(cdt)
    return ops


def nonlinear_kind(p: OracleParams) -> str:
    """Dispatch of ASS:1454-1492."""
    if not p.useSplitStep:
        return 'rusanov'
    if p.fluxScheme == 'kt':
        return 'kt'
    return 'tvd' if p.useTVD else 'rusanov'


def beta_tvd_for(p: OracleParams, dZ_step: float) -> float:
    """ASS:1465-1466."""
    norm_step = dZ_step / (p.c0 / p.f0)
    return max(1.0, 2.0 - p.adaptiveFilterStrength * norm_step * 10)


def march_step(field, ops, p: OracleParams, N, dZ, kt_subcycle=None, obl_map=None):
    """One composite step as dispatched at ASS:1454-1492; with ``obl_map`` (the
    precalculate_obliquity_map of ASS:1374-1377) the obliquity variants."""
    dT = p.dT
    kind = nonlinear_kind(p)
    if not p.useSplitStep:
        return march_step_sequential(field, ops['HH'], ops['abl'], ops['afilt'], N, dZ, dT, obl_map)
    if kind == 'kt':
        if kt_subcycle == 'column':
            nl = lambda u, n: kt_flux_subcycled_columns(u, n, dZ, dT)
        elif kt_subcycle == 'global':
            nl = lambda u, n: kt_flux_adaptive(u, n, dZ, dT)
        else:
            nl = lambda u, n: kt_flux(u, n, dZ, dT)
    elif kind == 'tvd':
        b = beta_tvd_for(p, dZ)
        nl = lambda u, n: rusanov_flux_tvd(u, n, dZ, dT, b)
    else:
        nl = lambda u, n: rusanov_flux_standard(u, n, dZ, dT)
    return march_step_split(field, ops['HH_half'], ops['abl_half'], ops['afilt_half'], N, dZ, dT, nl, obl_map)


def angular_spectrum_solve(initial_field, p: OracleParams, dtype=np.float64,
                           per_step_callback: Optional[Callable] = None,
                           max_steps_run: Optional[int] = None, kt_subcycle=None,
                           timings: Optional[dict] = None, tof_ref_trace=None, tof_t_ref_peak_s=None,
                           tof_env_ratio=None, taxis=None):
    """ASS:1217-1567 restated.

    ``max_steps_run`` stops after that many steps (used to time a bounded
    sample); ``timings`` receives 'setup_s' and per-step wall times 'step_s'.
    Returns (field, pnp, ppp, pI, pIloss, zaxis, pax) like ASS:1567, with a
    per-step TOF map appended (ASS:1387-1397, :1538-1545, :1563-1566) when
    ``tof_ref_trace`` or ``tof_env_ratio`` is given; the TOF estimators are
    ``tof_oracle``'s restatement of the absent ``tof_extraction`` module.
    """
    import time
    t_setup = time.perf_counter()
    dtype = np.dtype(dtype)
    init = np.asarray(initial_field)
    nX, nY, nT = init.shape
    c0 = p.c0
    N = p.beta / (2 * c0 ** 3 * p.rho0)
    max_amp = np.max(np.abs(init))
    dZ = min(p.stabilityRecoveryFactor * p.dT / (max_amp * N + 1e-30), p.dZmin) if max_amp > 0 else p.dZmin
    dZ = max(dZ, p.dZmin)
    ops = build_operators(nX, nY, nT, p, dZ, dtype)
    abl_freq = None
    if p.useFreqWeightedBoundary:
        abl_freq = precalculate_abl_freq(nX, nY, nT, p.dT, c0, p.f0, p.boundaryFactor,
                                         p.boundaryProfile, dtype)
    if p.useSuperAbsorbing:
        wx = make_boundary_weights(nX, boundary_width(nX, p.boundaryFactor))
        wy = make_boundary_weights(nY, boundary_width(nY, p.boundaryFactor))
    f0_bin = p.f0 / (1.0 / (nT * p.dT))
    obl_map = (precalculate_obliquity_map(nX, nY, nT, p.dX, p.dY, p.dT, c0)
               if p.useNonlinearityObliquity else None)
    tof_mf = tof_ref_trace is not None
    tof_on = tof_mf or tof_env_ratio is not None
    if tof_mf and tof_t_ref_peak_s is None:
        raise ValueError('tof_t_ref_peak_s is required when tof_ref_trace is provided')
    tof_t = np.asarray(taxis, dtype=np.float64) if taxis is not None else np.arange(nT, dtype=np.float64) * p.dT
    from oracle import tof_oracle

    def alloc(ms):
        return (np.zeros((nX, nY, ms), np.float32), np.zeros((nX, nY, ms), np.float32),
                np.zeros((nX, nY, ms), np.float32), np.zeros((nX, nY, ms), np.float32),
                np.zeros((nT, ms), np.float32))

    max_steps = int(np.ceil(p.propDist / dZ)) + 100
    pnp, ppp, pI, pIloss, pax = alloc(max_steps)
    tof = np.zeros((nX, nY, max_steps), np.float32) if tof_on else None
    field = init.astype(dtype)
    zvec = []
    cc = 0
    if timings is not None:
        timings['setup_s'] = time.perf_counter() - t_setup
        timings['step_s'] = []
    while sum(zvec) < p.propDist - 1e-15:
        if max_steps_run is not None and cc >= max_steps_run:
            break
        t_step = time.perf_counter()
        remaining = p.propDist - sum(zvec)
        dZ_step = min(dZ, remaining)
        max_field = np.max(np.abs(field))
        if N * dZ_step / p.dT * max_field > p.stabilityThreshold:
            new_dZ = max(p.stabilityRecoveryFactor * p.dT / (max_field * N), p.dZmin)
            if new_dZ == dZ:
                raise StabilityError('stability criterion violated and dZ cannot shrink below dZmin')
            dZ = new_dZ
            cc = 0
            zvec = []
            field = init.astype(dtype)
            ops = build_operators(nX, nY, nT, p, dZ, dtype)
            max_steps = int(np.ceil(p.propDist / dZ)) + 100
            pnp, ppp, pI, pIloss, pax = alloc(max_steps)
            tof = np.zeros((nX, nY, max_steps), np.float32) if tof_on else None
            continue
        I_before = np.sum(field.astype(np.float64) ** 2, axis=2)
        step_ops = ops
        if abs(dZ_step - dZ) > 1e-15:
            step_ops = build_operators(nX, nY, nT, p, dZ_step, dtype)
        field = march_step(field, step_ops, p, N, dZ_step, kt_subcycle, obl_map)
        if abl_freq is not None:
            field = freq_weighted_boundary_step(field, abl_freq)
        if p.useSuperAbsorbing:
            field = super_absorbing_step(field, wx, wy, p.dX, p.dY, p.dT, c0, p.superAbsorbingStrength)
        z_cur = sum(zvec)
        if p.phaseScreens is not None:
            for scr in p.phaseScreens:
                if z_cur < scr[0] <= z_cur + dZ_step:
                    amp = scr[2] if len(scr) > 2 else None
                    field = apply_phase_screen(field, np.asarray(scr[1]), f0_bin, amp)
        if p.sourcePlanes is not None:
            for sp_z, sp in p.sourcePlanes:
                if z_cur < sp_z <= z_cur + dZ_step:
                    field = (field + np.asarray(sp).astype(np.float32).astype(dtype)).astype(dtype)
        I_after = np.sum(field.astype(np.float64) ** 2, axis=2)
        pIloss[:, :, cc] = np.maximum(0, I_before - I_after)
        pI[:, :, cc] = I_after
        pnp[:, :, cc] = np.min(field, axis=2)
        ppp[:, :, cc] = np.max(field, axis=2)
        pax[:, cc] = field[nX // 2, nY // 2, :]
        zvec.append(dZ_step)
        if tof_on:
            if tof_mf:
                tof[:, :, cc] = tof_oracle.extract_tof_matched_filter_parabolic(
                    field, tof_t, sum(zvec), c0, tof_ref_trace, float(tof_t_ref_peak_s))
            else:
                tof[:, :, cc] = tof_oracle.extract_tof_envelope(field, tof_t, sum(zvec), c0, tof_env_ratio)
        if per_step_callback is not None:
            per_step_callback(cc, sum(zvec), field)
        cc += 1
        if timings is not None:
            timings['step_s'].append(time.perf_counter() - t_step)
    zaxis = np.cumsum(zvec[:cc])
    out = (field, pnp[:, :, :cc], ppp[:, :, :cc], pI[:, :, :cc], pIloss[:, :, :cc], zaxis, pax[:, :cc])
    return out + (tof[:, :, :cc],) if tof_on else out
