"""Synthetic workloads of the benchmark configurations (deterministic inputs).

Conventions follow the reference's own test inputs: centred axes
``(i - (n-1)/2) d`` and the Gaussian-enveloped tone burst
``exp(-(1.05 t w0/(ncyc pi))^(2 dur)) sin(w0 t) p0`` with dur = 2
(validate_solver.py:41-57, validate_analytical.py:160-174).
"""
from __future__ import annotations

import numpy as np

from .sources import make_bowl_source_planes, skull_screen

MEDIUM = dict(c0=1540.0, rho0=1000.0, beta=3.5, alpha0=0.5, attenPow=1.1)


def axes(n, d):
    return (np.arange(n) - (n - 1) / 2) * d


def tone_burst(t, f0, p0, ncycles, dur=2):
    w0 = 2 * np.pi * f0
    return np.exp(-(1.05 * t * w0 / (ncycles * np.pi)) ** (2 * dur)) * np.sin(w0 * t) * p0


def params_kwargs(**over):
    kw = dict(MEDIUM)
    kw.update(f0=1e6, fluxScheme='kt', useSplitStep=True, boundaryProfile='wendland')
    kw.update(over)
    return kw


def gaussian_beam(nX=64, nY=64, nT=256, dX=3.08e-4, dT=4e-8, f0=1e6, p0=2e5, ncycles=3, sigma=4e-3):
    """Small plain-method case: Gaussian aperture (sigma) times a tone burst."""
    x, y, t = axes(nX, dX), axes(nY, dX), axes(nT, dT)
    ap = np.exp(-(x[:, None] ** 2 + y[None, :] ** 2) / (2 * sigma ** 2))
    field = ap[:, :, None] * tone_burst(t, f0, p0, ncycles)[None, None, :]
    return field.astype(np.float32)


def config1(**over):
    """64x64x256, Gaussian aperture 4 mm, 1 MHz 3 cycles, 200 kPa, propDist 2 cm, dZ = dX."""
    dX = 3.08e-4
    field = gaussian_beam()
    kw = params_kwargs(dX=dX, dY=dX, dT=4e-8, propDist=0.02, dZmin=dX)
    kw.update(over)
    return field, kw


def focused_piston(nX=256, nY=256, nT=1024, dX=3.08e-4, dT=4e-8, f0=1e6, p0=1e6, ncycles=5,
                   radius=30e-3, focus=50e-3, c0=1540.0, steer=(0.0, 0.0)):
    """Disk aperture with a focusing delay law; the burst stays centred in the window.

    Delay tau(r) = -(|r - s|^2)/(2 F c0) + a^2/(4 F c0) with s the lateral
    steering offset of the focal point (zero for an on-axis focus).
    """
    x, y, t = axes(nX, dX), axes(nY, dX), axes(nT, dT)
    r2 = x[:, None] ** 2 + y[None, :] ** 2
    ap = (r2 <= radius ** 2).astype(np.float64)
    rs2 = (x[:, None] - steer[0]) ** 2 + (y[None, :] - steer[1]) ** 2
    tau = -rs2 / (2 * focus * c0) + radius ** 2 / (4 * focus * c0)
    field = ap[:, :, None] * tone_burst(t[None, None, :] - tau[:, :, None], f0, p0, ncycles)
    return field.astype(np.float32)


def config2(**over):
    """256x256x1024 focused piston (r 30 mm, F 50 mm), 1 MHz 5 cycles, 1 MPa, 6.5 cm,
    dZ = dX (212 steps), FW + super-absorbing boundaries, boundaryFactor 0.1."""
    dX = 3.08e-4
    field = focused_piston()
    kw = params_kwargs(dX=dX, dY=dX, dT=4e-8, propDist=0.065, dZmin=dX, boundaryFactor=0.1,
                       useFreqWeightedBoundary=True, useSuperAbsorbing=True)
    kw.update(over)
    return field, kw


def config3(**over):
    """384x384x1024 bowl (r 46 mm, roc 80 mm, focus 50 mm, inner 20.5 mm, 8 elements,
    400 kPa) injected plane by plane, skull screen at z = 10 mm."""
    nX = nY = 384
    nT = 1024
    dX = 3.08e-4
    dT = 4e-8
    f0 = 1e6
    x, y, t = axes(nX, dX), axes(nY, dX), axes(nT, dT)
    planes, depth = make_bowl_source_planes(x, y, t, f0, MEDIUM['c0'], 4e5, 46e-3, 80e-3, dX,
                                            ncycles=3, dur=2, inner_radius=20.5e-3, focus=50e-3,
                                            n_elements=8)
    phase, amp, _ = skull_screen(nX, nY, dX, dX, f0, seed=3)
    kw = params_kwargs(dX=dX, dY=dX, dT=dT, propDist=0.065, dZmin=dX, boundaryFactor=0.1,
                       phaseScreens=[(10e-3, phase, amp)], sourcePlanes=planes[1:])
    kw.update(over)
    return planes[0][1], kw


def config4(n=32, seed=0, propDist=0.02, sizes=(128, 192, 256, 384), lengths=(512, 1024, 2048)):
    """Batched transmit configurations: seeded focused pistons of uneven size.

    Instance k draws nX = nY from ``sizes`` and nT from ``lengths``, a focus
    depth uniform in 30-70 mm, a lateral (x) steering of the focal point
    uniform in +-10 mm and p0 uniform in 0.2-1.5 MPa; 1 MHz, 5 cycles, dZ = dX,
    Wendland and frequency-weighted boundaries (boundaryFactor 0.1).  The
    super-absorbing pass is left off: at nT = 2048 on 384^2 grids its
    low-frequency coefficient c0/omega (ASS:225-251) makes the reference's march
    grow without bound within ~45 steps (the FP64 restatement does the same).
    The piston radius is
    0.3 nX dX capped at focus / 2.5: at f-number >= 1.25 the linear focal gain
    k a^2 / (2 F) stays below ~36, so the focal pressure stays under the
    solver's stability limit at dZ = dX (0.2 dT / (N dZ) ~ 54 MPa; above it the
    solve raises StabilityError, ASS:1417-1439).
    Returns [(field, kwargs, meta)] with meta the drawn values."""
    rng = np.random.default_rng(seed)
    dX, dT = 3.08e-4, 4e-8
    out = []
    for _ in range(n):
        nx = int(rng.choice(sizes))
        nt = int(rng.choice(lengths))
        focus = float(rng.uniform(30e-3, 70e-3))
        steer = float(rng.uniform(-10e-3, 10e-3))
        p0 = float(rng.uniform(0.2e6, 1.5e6))
        radius = min(0.3 * nx * dX, focus / 2.5)
        field = focused_piston(nX=nx, nY=nx, nT=nt, dX=dX, dT=dT, p0=p0, radius=radius, focus=focus,
                               steer=(steer, 0.0))
        kw = params_kwargs(dX=dX, dY=dX, dT=dT, propDist=propDist, dZmin=dX, boundaryFactor=0.1,
                           useFreqWeightedBoundary=True)
        out.append((field, kw, dict(nX=nx, nT=nt, focus=focus, steer=steer, p0=p0, radius=radius)))
    return out


def voxels(shape):
    return int(np.prod(shape))
