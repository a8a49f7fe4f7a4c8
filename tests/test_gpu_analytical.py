"""The reference's focused-piston configuration sweep (validate_analytical.py:464-673,
test C2) through the HIP march, against its published accuracy.

A 3 mm (8 wavelength) piston focused at F = 20 mm, 4 MHz in water (alpha0 = -1,
beta = 0), 193 x 193 x 1201 samples (general-size kernels on every axis, odd nT),
marched 26 mm at dZ = 2 dX and 8 dX, in the four solver configurations
{sequential, split-step} x {no k-filter, adaptive k-filter}.  The metrics are the
reference's: the RMS difference of the normalised on-axis peak amplitude from the
paraxial O'Neil solution, and of the normalised focal-plane intensity from the
Airy pattern (2 J1(u)/u)^2.  The reference publishes them in the legends of
validation_results/analytical/testC2_focused_piston_config_sweep.png (z spacing
0.6 mm = 8 dX: 0.088 / 0.003 for all four configurations) and
testC2_focused_piston_config_sweep_fine.png (0.15 mm = 2 dX: 0.091 / 0.003).
The legends round to three decimals; the bound adds a margin of 1e-3 for the
single-precision march.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F0, C0 = 4e6, 1500.0
LAM = C0 / F0
DX = LAM / 5
DT = DX / (5 * C0)
NX = NY = 193
NT = 1201
A = 8 * LAM
F_FOC = 20e-3
P0 = 0.3e6


def _initial_field(ncycles=4.0):
    """The C2 source (validate_analytical.py:511-544): Gaussian-windowed tone burst,
    rolled by the focusing delay rounded to whole samples, on the disk aperture."""
    x = np.arange(NX) * DX - (NX - 1) * DX / 2
    y = np.arange(NY) * DX - (NY - 1) * DX / 2
    t = np.arange(NT) * DT - (NT - 1) * DT / 2
    r2 = x[:, None] ** 2 + y[None, :] ** 2
    w0 = 2 * np.pi * F0
    burst = (np.exp(-(1.05 * t * w0 / (ncycles * np.pi)) ** 4) * np.sin(t * w0) * P0).astype(np.float32)
    shift = np.round((F_FOC - np.sqrt(F_FOC ** 2 + r2)) / C0 / DT).astype(np.int64)
    field = np.zeros((NX, NY, NT), dtype=np.float32)
    for ix, iy in zip(*np.nonzero(r2 <= A ** 2)):
        field[ix, iy] = np.roll(burst, shift[ix, iy])
    return field, x


def _metrics(pI, zaxis, pax, x):
    """On-axis and focal-plane RMS of validate_analytical.py:585-609."""
    k0 = 2 * np.pi * F0 / C0
    onaxis = np.abs(pax).max(axis=0)
    iF = int(np.argmin(np.abs(zaxis - F_FOC)))
    lat = pI[:, NY // 2, iF].astype(np.float64)
    lat = lat / (lat.max() + 1e-30)
    zs = np.maximum(zaxis, 1e-9)
    den = F_FOC - zaxis
    ana = 2 * P0 * np.abs(F_FOC / np.where(np.abs(den) < 1e-9, 1e-9, den)) * \
        np.abs(np.sin(k0 * A ** 2 * (F_FOC - zaxis) / (4 * F_FOC * zs)))
    near = np.abs(zaxis - F_FOC) < 2 * (zaxis[1] - zaxis[0])
    ana[near] = P0 * np.pi * A ** 2 / (LAM * F_FOC)
    from scipy.special import j1
    u = np.maximum(k0 * A * np.abs(x) / F_FOC, 1e-10)
    airy = (2 * j1(u) / u) ** 2
    airy[k0 * A * np.abs(x) / F_FOC < 1e-10] = 1.0
    airy = airy / airy.max()
    on_rms = float(np.sqrt(np.mean((onaxis / onaxis.max() - ana / ana.max()) ** 2)))
    lat_rms = float(np.sqrt(np.mean((lat - airy) ** 2)))
    return on_rms, lat_rms


CONFIGS = [("sequential, no filter", False, False), ("split-step, no filter", True, False),
           ("sequential, k-filter", False, True), ("split-step, k-filter", True, True)]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("dz_factor,pub_onaxis,pub_lateral", [(8, 0.088, 0.003), (2, 0.091, 0.003)])
def test_focused_piston_sweep_reproduces_published_rms(dz_factor, pub_onaxis, pub_lateral):
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field0, x = _initial_field()
    configs = CONFIGS if dz_factor == 8 else CONFIGS[::3]
    for label, split, kfilter in configs:
        p = SolverParams(dX=DX, dY=DX, dT=DT, c0=C0, rho0=1000, beta=0.0, alpha0=-1, f0=F0,
                         propDist=1.3 * F_FOC, boundaryFactor=0.15, useBoundaryLayer=True, useTVD=False,
                         stabilityThreshold=0.2, stabilityRecoveryFactor=0.15, dZmin=DX * dz_factor,
                         useSplitStep=split, useAdaptiveFiltering=kfilter)
        out = angular_spectrum_solve(field0, p, verbose=False)
        pI, zaxis, pax = np.asarray(out[3]), np.asarray(out[5]), np.asarray(out[6])
        assert len(zaxis) == int(np.ceil(1.3 * F_FOC / (DX * dz_factor) - 1e-9))
        on_rms, lat_rms = _metrics(pI, zaxis, pax, x)
        print(f"dZ = {dz_factor} dX, {label}: on-axis RMS {on_rms:.4f} (published {pub_onaxis}), "
              f"focal-plane RMS {lat_rms:.4f} (published {pub_lateral})", flush=True)
        assert abs(on_rms - pub_onaxis) <= 1.5e-3
        assert abs(lat_rms - pub_lateral) <= 1.5e-3
