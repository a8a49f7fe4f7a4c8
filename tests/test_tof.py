"""CPU pins of the TOF oracle (oracle/tof_oracle.py).  The reference's module is not
shipped (angular_spectrum_solver.py:26-29), so parity against it is unpinned; these
checks fix the restated estimators by closed forms."""
import numpy as np
import pytest

from oracle import tof_oracle as T

NT, DT, C0 = 512, 4e-8, 1540.0
TAX = (np.arange(NT) - (NT - 1) / 2) * DT


def _burst(delay, sigma, f0=3e6, amp=1.0):
    tm = TAX - delay
    return amp * np.exp(-tm ** 2 / (2 * sigma ** 2)) * np.cos(2 * np.pi * f0 * tm)


def test_envelope_peak_sits_at_the_delay_plus_z_over_c0():
    delays = TAX[[180, 256, 301]]                       # on the sample grid
    f = np.stack([_burst(d, 20 * DT) for d in delays])[None]   # (1, 3, NT)
    z = 0.013
    tof = T.extract_tof_envelope(f, TAX, z, C0, 1.0)
    np.testing.assert_allclose(tof[0], delays + z / C0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("ratio", [0.5, 0.1])
def test_envelope_threshold_crossing_of_a_gaussian(ratio):
    sigma = 25 * DT
    d = TAX[250] + 0.37 * DT
    f = _burst(d, sigma)[None, None, :]
    tof = T.extract_tof_envelope(f, TAX, 0.0, C0, ratio)[0, 0]
    # the Gaussian envelope reaches ratio of its peak sigma sqrt(-2 ln r) before the peak;
    # linear interpolation between samples leaves O(dt^2) of it
    want = d - sigma * np.sqrt(-2 * np.log(ratio))
    assert abs(tof - want) < 0.02 * DT, (tof - want) / DT


def test_analytic_signal_convention():
    """Real part is the signal; the imaginary part of cos is sin (away from DC and Nyquist)."""
    n = 64
    t = np.arange(n)
    for m in (1, 5, 31):
        a = T.analytic_signal(np.cos(2 * np.pi * m * t / n))
        np.testing.assert_allclose(a.real, np.cos(2 * np.pi * m * t / n), atol=1e-12)
        np.testing.assert_allclose(a.imag, np.sin(2 * np.pi * m * t / n), atol=1e-12)
    a = T.analytic_signal(np.cos(np.pi * t))         # Nyquist: kept once, no quadrature part
    np.testing.assert_allclose(a, np.cos(np.pi * t), atol=1e-12)


def _shifted(ref, shift):
    """ref delayed by `shift` samples (circular, band-limited)."""
    F = np.fft.rfft(ref)
    k = np.arange(len(F))
    return np.fft.irfft(F * np.exp(-2j * np.pi * k * shift / NT), n=NT)


@pytest.mark.parametrize("shift", [0, 7, -12, 40])
def test_matched_filter_integer_shift_is_exact(shift):
    ref = _burst(TAX[200], 15 * DT)
    f = _shifted(ref, shift)[None, None, :]
    t_ref_peak = TAX[200]
    tof = T.extract_tof_matched_filter_parabolic(f, TAX, 0.02, C0, ref, t_ref_peak)[0, 0]
    # the correlation is symmetric about an integer lag: the parabola's vertex is the sample
    assert abs(tof - (t_ref_peak + shift * DT + 0.02 / C0)) < 1e-6 * DT


@pytest.mark.parametrize("shift", [0.3, -2.45, 17.8])
def test_matched_filter_subsample_shift(shift):
    ref = _burst(TAX[260], 15 * DT, f0=1.5e6)
    f = _shifted(ref, shift)[None, None, :]
    tof = T.extract_tof_matched_filter_parabolic(f, TAX, 0.0, C0, ref, TAX[260])[0, 0]
    assert abs(tof - (TAX[260] + shift * DT)) < 0.05 * DT


def test_parabolic_vertex_is_exact_on_a_parabola():
    """With a reference that is a unit impulse the correlation is the trace itself, so a
    trace that is a sampled parabola (peak off the grid) has its vertex recovered exactly."""
    ref = np.zeros(NT)
    ref[0] = 1.0
    v = 100.3
    k = np.arange(NT)
    trace = 5.0 - (k - v) ** 2
    trace = np.where(np.abs(k - v) < 3, trace, -50.0)
    tof = T.extract_tof_matched_filter_parabolic(trace[None, None, :], TAX, 0.0, C0, ref, 0.0)[0, 0]
    assert abs(tof - v * DT) < 1e-9 * DT


def test_zero_column_has_no_arrival():
    f = np.zeros((1, 2, NT))
    f[0, 1] = _burst(TAX[256], 20 * DT)
    e = T.extract_tof_envelope(f, TAX, 0.0, C0, 0.5)
    m = T.extract_tof_matched_filter_parabolic(f, TAX, 0.0, C0, f[0, 1], TAX[256])
    assert np.isnan(e[0, 0]) and np.isfinite(e[0, 1])
    assert np.isnan(m[0, 0]) and np.isfinite(m[0, 1])
