"""CPU oracle of the time-of-flight outputs of the march.

TEST INFRASTRUCTURE ONLY: like ``asm_oracle``, only ``tests/`` (and the timed CPU
baseline) may use it, and only as the checker.

The reference imports ``extract_tof_envelope`` and
``extract_tof_matched_filter_parabolic`` from a module ``tof_extraction``
(angular_spectrum_solver.py:26-29) that it does not ship; its solver calls them
once per step on the field after the step (ASS:1538-1545) with the time axis,
the cumulative z and c0, and documents them only in its docstring
(ASS:1244-1264).  PARITY UNPINNED against the reference: the functions below
restate that docstring as follows, and are pinned in ``tests/test_tof.py`` by
closed forms (the envelope peak of a delayed tone burst sits at its delay; a
Gaussian envelope crosses r of its maximum sigma sqrt(-2 ln r) before the peak;
parabolic interpolation is exact on a sampled parabola; a sub-sample shift of
the reference is recovered):

* The march's time axis is retarded time (the pulse stays in the window as it
  travels), so an arrival time at depth z is the retarded time plus z / c0.
* Envelope: the magnitude of the analytic signal (scipy.signal.hilbert's
  convention: bins 1..ceil(nT/2)-1 doubled, the Nyquist bin of even nT kept
  once, negative frequencies removed).  TOF is the first time the envelope
  reaches env_ratio times its maximum, interpolated linearly between that
  sample and the one before (env_ratio = 1 gives the envelope-peak sample).
* Matched filter: the circular cross-correlation c[k] = sum_t p[t] ref[t - k]
  (computed in the rfft domain); k* is its first maximum, refined by the vertex
  of the parabola through c[k*-1], c[k*], c[k*+1] (circular neighbours); the
  lag k* + delta is taken in (-nT/2, nT/2] samples and TOF = t_ref_peak + lag dt.
* A column that is zero everywhere has no arrival: NaN.
"""
from __future__ import annotations

import numpy as np


def analytic_signal(p: np.ndarray) -> np.ndarray:
    """Analytic signal along the last axis (scipy.signal.hilbert's convention)."""
    n = p.shape[-1]
    F = np.fft.fft(p, axis=-1)
    h = np.zeros(n)
    h[0] = 1.0
    if n % 2 == 0:
        h[n // 2] = 1.0
        h[1:n // 2] = 2.0
    else:
        h[1:(n + 1) // 2] = 2.0
    return np.fft.ifft(F * h, axis=-1)


def extract_tof_envelope(field, taxis, z, c0, env_ratio):
    """(nX, nY) arrival times: envelope threshold crossing (retarded) + z / c0."""
    p = np.asarray(field, dtype=np.float64)
    t = np.asarray(taxis, dtype=np.float64)
    env = np.abs(analytic_signal(p))
    emax = env.max(axis=-1, keepdims=True)
    thr = env_ratio * emax
    above = env >= thr
    k = np.argmax(above, axis=-1)
    kk = k[..., None]
    e1 = np.take_along_axis(env, kk, axis=-1)[..., 0]
    e0 = np.take_along_axis(env, np.maximum(kk - 1, 0), axis=-1)[..., 0]
    th = thr[..., 0]
    denom = e1 - e0
    frac = np.where((k > 0) & (denom > 0), (th - e0) / np.where(denom > 0, denom, 1.0), 1.0)
    t1 = t[k]
    t0 = t[np.maximum(k - 1, 0)]
    tof = np.where(k > 0, t0 + frac * (t1 - t0), t1)
    tof = np.where(emax[..., 0] > 0, tof, np.nan)
    return tof + z / c0


def extract_tof_matched_filter_parabolic(field, taxis, z, c0, ref_trace, t_ref_peak):
    """(nX, nY) arrival times: t_ref_peak + the sub-sample lag of the correlation peak + z / c0."""
    p = np.asarray(field, dtype=np.float64)
    t = np.asarray(taxis, dtype=np.float64)
    n = p.shape[-1]
    dt = t[1] - t[0]
    R = np.fft.rfft(np.asarray(ref_trace, dtype=np.float64))
    c = np.fft.irfft(np.fft.rfft(p, axis=-1) * np.conj(R), n=n, axis=-1)
    k = np.argmax(c, axis=-1)
    kk = k[..., None]
    cm = np.take_along_axis(c, (kk - 1) % n, axis=-1)[..., 0]
    c0v = np.take_along_axis(c, kk, axis=-1)[..., 0]
    cp = np.take_along_axis(c, (kk + 1) % n, axis=-1)[..., 0]
    den = cm - 2.0 * c0v + cp
    delta = np.where(den < 0, 0.5 * (cm - cp) / np.where(den < 0, den, -1.0), 0.0)
    lag = k + delta
    lag = np.where(lag > n / 2, lag - n, lag)
    tof = t_ref_peak + lag * dt
    tof = np.where(np.abs(p).max(axis=-1) > 0, tof, np.nan)
    return tof + z / c0
