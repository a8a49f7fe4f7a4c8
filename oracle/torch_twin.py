"""Torch FP64 twin of the CPU oracle's driver (TEST INFRASTRUCTURE ONLY).

Same step functions as ``asm_oracle`` (each restating the reference function
cited there), evaluated with ``torch.fft`` in float64/complex128 on any torch
device.  Operators come from the NumPy oracle's precompute, so the twin only
restates the per-step arithmetic.  It exists to give full-size FP64 references
(e.g. 256x256x1024 over 212 steps) on the GPU host, where the NumPy oracle would
take hours; it is pinned against the NumPy oracle at small sizes by the CPU test
suite (agreement ~1e-12).
"""
from __future__ import annotations

import numpy as np
import torch

from . import asm_oracle as O


def _fftshift(x, dims):
    return torch.fft.fftshift(x, dim=dims)


def angular_spectrum_step(field, HH, abl):
    """ASS:649-655."""
    F = torch.fft.fftn(field)
    F = torch.fft.ifftshift(torch.fft.fftshift(F) * HH)
    return torch.fft.ifftn(F).real * abl


def attenuation_step(field, afilt):
    """ASS:667-685."""
    nT = field.shape[2]
    F = torch.fft.rfft(field, dim=2)
    F = torch.fft.fftshift(torch.fft.fft2(F, dim=(0, 1)), dim=(0, 1)) * afilt
    F = torch.fft.ifft2(torch.fft.ifftshift(F, dim=(0, 1)), dim=(0, 1))
    return torch.fft.irfft(F, n=nT, dim=2)


def freq_weighted_boundary_step(field, abl_freq):
    """ASS:658-664."""
    nT = field.shape[2]
    return torch.fft.irfft(torch.fft.rfft(field, dim=2) * abl_freq, n=nT, dim=2)


def super_absorbing_step(field, wx, wy, dX, dY, dT, c0, strength):
    """ASS:185-272."""
    nX, nY, nT = field.shape
    F = torch.fft.rfft(field, dim=2)
    nF = F.shape[2]
    freqs = np.fft.rfftfreq(nT, dT)
    omega = 2.0 * np.pi * freqs
    omega_min = 2.0 * np.pi * freqs[1] if nF > 1 else 1.0
    valid = np.abs(omega) > 0.5 * omega_min
    coeff = np.where(valid, -1j * c0 / np.where(valid, omega, 1.0), 0.0 + 0j)
    coeff = torch.from_numpy(coeff).to(F.device)[None, None, :]
    dF = torch.empty_like(F)
    dF[1:-1] = (F[2:] - F[:-2]) / (2 * dX)
    dF[0] = (F[1] - F[0]) / dX
    dF[-1] = (F[-1] - F[-2]) / dX
    ix = torch.arange(nX, device=F.device)
    wxt = torch.as_tensor(np.asarray(wx, dtype=np.float64), device=F.device)
    wl = (wxt * (ix < nX // 2))[:, None, None]
    wr = (wxt * (ix >= nX // 2))[:, None, None]
    F = F - strength * wl * (0.5 * (F + coeff * dF))
    F = F - strength * wr * (0.5 * (F - coeff * dF))
    dF = torch.empty_like(F)
    dF[:, 1:-1] = (F[:, 2:] - F[:, :-2]) / (2 * dY)
    dF[:, 0] = (F[:, 1] - F[:, 0]) / dY
    dF[:, -1] = (F[:, -1] - F[:, -2]) / dY
    iy = torch.arange(nY, device=F.device)
    wyt = torch.as_tensor(np.asarray(wy, dtype=np.float64), device=F.device)
    wb = (wyt * (iy < nY // 2))[None, :, None]
    wt = (wyt * (iy >= nY // 2))[None, :, None]
    F = F - strength * wb * (0.5 * (F + coeff * dF))
    F = F - strength * wt * (0.5 * (F - coeff * dF))
    return torch.fft.irfft(F, n=nT, dim=2)


def apply_phase_screen(field, phase, f0_bin, amp=None):
    """ASS:352-390."""
    nT = field.shape[2]
    F = torch.fft.rfft(field, dim=2)
    nF = F.shape[2]
    fs = (torch.arange(nF, dtype=torch.float64, device=F.device) / max(f0_bin, 1.0))[None, None, :]
    F = F * torch.exp(-1j * (phase[:, :, None] * fs))
    if amp is not None:
        F = F * torch.pow(torch.clamp(amp, 1e-6, 1.0)[:, :, None], fs)
    return torch.fft.irfft(F, n=nT, dim=2)


def kt_rhs(u, N, dT, limiter='mc'):
    """ASS:758-797 / :800-830."""
    dm = u[..., 1:-1] - u[..., :-2]
    dp = u[..., 2:] - u[..., 1:-1]
    if limiter == 'mc':
        a, b, c = 2 * dm, 0.5 * (dm + dp), 2 * dp
        same = (a * b > 0) & (b * c > 0)
        m = torch.minimum(torch.minimum(a.abs(), b.abs()), c.abs())
        sm = torch.where(same, torch.sign(a) * m, torch.zeros_like(a))
    else:
        sm = torch.where(dm * dp > 0, torch.sign(dm) * torch.minimum(dm.abs(), dp.abs()), torch.zeros_like(dm))
    sigma = torch.zeros_like(u)
    sigma[..., 1:-1] = sm
    um = u[..., :-1] + 0.5 * sigma[..., :-1]
    up = u[..., 1:] - 0.5 * sigma[..., 1:]
    ap = torch.clamp(torch.maximum(-um, -up), min=0.0)
    am = torch.clamp(torch.minimum(-um, -up), max=0.0)
    fm = -0.5 * um ** 2
    fp = -0.5 * up ** 2
    den = ap - am
    small = den.abs() < 1e-14
    sden = torch.where(small, torch.ones_like(den), den)
    kt = (ap * fm - am * fp - ap * am * (up - um)) / sden
    lf = 0.5 * (fm + fp) - 0.5 * torch.maximum(um.abs(), up.abs()) * (up - um)
    fh = torch.where(small, lf, kt)
    rhs = torch.zeros_like(u)
    rhs[..., 1:-1] = -(N / dT) * (fh[..., 1:] - fh[..., :-1])
    return rhs


def kt_flux(u, N, dZ, dT):
    """ASS:833-839."""
    r0 = kt_rhs(u, N, dT)
    s1 = u + dZ * r0
    return 0.5 * u + 0.5 * (s1 + dZ * kt_rhs(s1, N, dT))


def rusanov_flux_standard(u, N, dZ, dT):
    """ASS:688-696."""
    lam = torch.maximum(u[..., :-1].abs(), u[..., 1:].abs())
    fh = -(u[..., :-1] ** 2 + u[..., 1:] ** 2) / 2 - lam * (u[..., 1:] - u[..., :-1])
    out = u.clone()
    out[..., 1:-1] += -(N * dZ / dT) * (fh[..., 1:] - fh[..., :-1])
    return out


def rusanov_flux_tvd(u, N, dZ, dT, beta):
    """ASS:699-724."""
    uL, uR = u[..., :-1], u[..., 1:]
    a = torch.maximum(uL.abs(), uR.abs())
    fh = 0.5 * (-0.5 * (uL ** 2 + uR ** 2)) - 0.5 * a * (uR - uL)
    fd = fh[..., 1:] - fh[..., :-1]
    dm = u[..., 1:-1] - u[..., :-2]
    dp = u[..., 2:] - u[..., 1:-1]
    eps = 1e-30
    r = dp / (dm + eps * torch.sign(dm + eps))
    ri = dm / (dp + eps * torch.sign(dp + eps))
    one = torch.ones_like(r)
    pp = torch.clamp(torch.minimum(torch.minimum(beta * r, one), torch.minimum(r, beta * one)), min=0.0)
    pm = torch.clamp(torch.minimum(torch.minimum(beta * ri, one), torch.minimum(ri, beta * one)), min=0.0)
    out = u.clone()
    out[..., 1:-1] += -(N * dZ / dT) * (0.5 * (pp + pm) * fd)
    return out


def _ops_to(ops, device):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in ops.items()}


def march_step(field, ops, p, N, dZ):
    """Dispatch of ASS:1454-1492 (obliquity variants excluded)."""
    dT = p.dT
    kind = O.nonlinear_kind(p)
    if not p.useSplitStep:
        field = angular_spectrum_step(field, ops['HH'], ops['abl'])
        field = rusanov_flux_standard(field, N, dZ, dT)
        return attenuation_step(field, ops['afilt'])
    field = attenuation_step(field, ops['afilt_half'])
    field = angular_spectrum_step(field, ops['HH_half'], ops['abl_half'])
    if kind == 'kt':
        field = kt_flux(field, N, dZ, dT)
    elif kind == 'tvd':
        field = rusanov_flux_tvd(field, N, dZ, dT, O.beta_tvd_for(p, dZ))
    else:
        field = rusanov_flux_standard(field, N, dZ, dT)
    field = angular_spectrum_step(field, ops['HH_half'], ops['abl_half'])
    return attenuation_step(field, ops['afilt_half'])


def source_slice_field(sp, device, dtype):
    """Dense field of a mask + delay source slice (the tone burst of ASS:1193-1202),
    synthesised on ``device`` in float64 and rounded to float32 as the reference
    stores it."""
    mask = torch.from_numpy(np.asarray(sp.mask)).to(device)
    d = torch.from_numpy(np.asarray(sp.delays, dtype=np.float64)).to(device)
    t = torch.from_numpy(np.asarray(sp.taxis, dtype=np.float64)).to(device)
    w0 = 2 * np.pi * sp.f0
    tm = t[None, None, :] - d[:, :, None]
    env = torch.exp(-(1.05 * tm * w0 / (sp.ncycles * np.pi)) ** (2 * sp.dur))
    fld = mask.to(torch.float64)[:, :, None] * env * torch.sin(w0 * tm) * sp.p0
    return fld.to(torch.float32).to(dtype)


def angular_spectrum_solve(initial_field, p, device='cpu', max_steps_run=None, dtype=np.float64,
                           progress=0):
    """Driver of ASS:1217-1567 on ``device`` (same logic as the NumPy oracle).

    dtype float64: the FP64 reference; float32: the reference's own dtype flow
    (operators built in FP64 and cast, ASS:1343-1351).  ``progress`` > 0 prints
    a line every that many steps.
    """
    init = np.asarray(initial_field)
    nX, nY, nT = init.shape
    c0 = p.c0
    N = p.beta / (2 * c0 ** 3 * p.rho0)
    max_amp = np.max(np.abs(init))
    dZ = min(p.stabilityRecoveryFactor * p.dT / (max_amp * N + 1e-30), p.dZmin) if max_amp > 0 else p.dZmin
    dZ = max(dZ, p.dZmin)
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    ops = _ops_to(O.build_operators(nX, nY, nT, p, dZ, dtype), device)
    abl_freq = None
    if p.useFreqWeightedBoundary:
        abl_freq = torch.from_numpy(O.precalculate_abl_freq(nX, nY, nT, p.dT, c0, p.f0, p.boundaryFactor,
                                                            p.boundaryProfile, dtype)).to(device)
    if p.useSuperAbsorbing:
        wx = O.make_boundary_weights(nX, O.boundary_width(nX, p.boundaryFactor)).astype(np.float64)
        wy = O.make_boundary_weights(nY, O.boundary_width(nY, p.boundaryFactor)).astype(np.float64)
    f0_bin = p.f0 / (1.0 / (nT * p.dT))
    max_steps = int(np.ceil(p.propDist / dZ)) + 100
    outs = torch.zeros((4, max_steps, nX, nY), dtype=torch.float64, device=device)
    pax = torch.zeros((max_steps, nT), dtype=torch.float64, device=device)
    field = torch.from_numpy(init.astype(np.float32).astype(dtype)).to(device)
    zvec = []
    cc = 0
    while sum(zvec) < p.propDist - 1e-15:
        if max_steps_run is not None and cc >= max_steps_run:
            break
        remaining = p.propDist - sum(zvec)
        dZ_step = min(dZ, remaining)
        max_field = float(field.abs().max())
        if N * dZ_step / p.dT * max_field > p.stabilityThreshold:
            raise O.StabilityError('stability criterion violated')
        I_before = (field.double() ** 2).sum(dim=2)
        step_ops = ops
        if abs(dZ_step - dZ) > 1e-15:
            step_ops = _ops_to(O.build_operators(nX, nY, nT, p, dZ_step, dtype), device)
        field = march_step(field, step_ops, p, N, dZ_step)
        if abl_freq is not None:
            field = freq_weighted_boundary_step(field, abl_freq)
        if p.useSuperAbsorbing:
            field = super_absorbing_step(field, wx, wy, p.dX, p.dY, p.dT, c0, p.superAbsorbingStrength)
        z_cur = sum(zvec)
        if p.phaseScreens is not None:
            for scr in p.phaseScreens:
                if z_cur < scr[0] <= z_cur + dZ_step:
                    ph = torch.from_numpy(np.asarray(scr[1]).astype(dtype)).to(device)
                    amp = None
                    if len(scr) > 2 and scr[2] is not None:
                        amp = torch.from_numpy(np.asarray(scr[2]).astype(dtype)).to(device)
                    field = apply_phase_screen(field, ph, f0_bin, amp)
        if p.sourcePlanes is not None:
            for sp_z, sp in p.sourcePlanes:
                if z_cur < sp_z <= z_cur + dZ_step:
                    if hasattr(sp, 'mask') and hasattr(sp, 'delays'):
                        field = field + source_slice_field(sp, device, tdt)
                    else:
                        field = field + torch.from_numpy(np.asarray(sp).astype(np.float32).astype(dtype)).to(device)
        I_after = (field.double() ** 2).sum(dim=2)
        outs[3, cc] = torch.clamp(I_before - I_after, min=0.0)
        outs[2, cc] = I_after
        outs[0, cc] = field.min(dim=2).values.double()
        outs[1, cc] = field.max(dim=2).values.double()
        pax[cc] = field[nX // 2, nY // 2, :].double()
        zvec.append(dZ_step)
        cc += 1
        if progress and cc % progress == 0:
            print(f'  twin step {cc} z={sum(zvec):.5f}', flush=True)
    res = outs[:, :cc].permute(0, 2, 3, 1).cpu().numpy()
    return (field.double().cpu().numpy(), res[0], res[1], res[2], res[3], np.cumsum(zvec[:cc]), pax[:cc].t().cpu().numpy())
