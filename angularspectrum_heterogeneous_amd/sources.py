"""Input builders: bowl sources (ASS:971-1211) and random phase screens (ASS:290-349).

``make_bowl_source_planes`` keeps the reference signature and return value
``(list[(z_mid, slice)], bowl_depth)``.  Each slice is a ``SourceSlice``: an
array-like object that materialises the reference's dense float32
(nX, nY, nT) field on ``np.asarray`` (so ``planes[0][1]`` works as an initial
field), while the solver injects it from its sparse form (a 2-D mask and delay
map) and synthesises the tone burst on the GPU.  At 384x384x1024 a dense slice
is 576 MiB; the sparse form is 1.3 MiB.  ``make_bowl_source`` (the one-plane
bowl) returns the dense field like the reference, or its ``SourceSlice`` with
``sparse=True``; the solver synthesises a ``SourceSlice`` initial field on the
GPU (``ash_source_field``).
"""
from __future__ import annotations

import math

import numpy as np


def _tone_burst(t_minus_d, omega0, p0, ncycles, dur):
    env = np.exp(-(1.05 * t_minus_d * omega0 / (ncycles * np.pi)) ** (2 * dur))
    return env, np.sin(omega0 * t_minus_d)


class SourceSlice:
    """One z-slice of a bowl source: mask, per-pixel delay and pulse parameters."""

    def __init__(self, mask, delays, taxis, f0, p0, ncycles, dur):
        self.mask = np.asarray(mask, dtype=bool)
        self.delays = np.asarray(delays, dtype=np.float64)
        self.taxis = np.asarray(taxis, dtype=np.float64)
        self.f0 = float(f0)
        self.p0 = float(p0)
        self.ncycles = float(ncycles)
        self.dur = dur
        self.shape = self.mask.shape + (self.taxis.shape[0],)
        self.dtype = np.dtype(np.float32)
        self.ndim = 3

    def dense(self) -> np.ndarray:
        """The reference's slice field (ASS:1193-1202), float32 (nX, nY, nT)."""
        omega0 = 2 * np.pi * self.f0
        t_grid = self.taxis[np.newaxis, np.newaxis, :]
        d_grid = self.delays[:, :, np.newaxis]
        env, sn = _tone_burst(t_grid - d_grid, omega0, self.p0, self.ncycles, self.dur)
        field = self.mask.astype(np.float32)[:, :, np.newaxis] * env * sn * self.p0
        return field.astype(np.float32)

    def __array__(self, dtype=None, copy=None):
        a = self.dense()
        return a if dtype is None else a.astype(dtype)

    def astype(self, dtype):
        return self.dense().astype(dtype)

    def device_field(self, device="cuda"):
        """The dense slice synthesised on the GPU (``ash_source_field``), float32 (nX, nY, nT)."""
        import torch

        from . import _lib
        lib = _lib.load()
        mask, delay, tax = self.device_arrays(device)
        out = torch.empty(self.shape, dtype=torch.float32, device=device)
        ncols = self.shape[0] * self.shape[1]
        _lib.check(lib.ash_source_field(mask.data_ptr(), delay.data_ptr(), tax.data_ptr(), ncols, self.shape[2],
                                        self.f0, self.p0, self.ncycles, float(self.dur), out.data_ptr(),
                                        _lib.current_stream()), "ash_source_field")
        torch.cuda.current_stream().synchronize()  # mask/delay/taxis are freed on return
        return out

    def device_arrays(self, device):
        import torch
        mask = torch.from_numpy(np.ascontiguousarray(self.mask.astype(np.uint8))).to(device)
        delay = torch.from_numpy(np.ascontiguousarray(self.delays)).to(device)
        tax = torch.from_numpy(np.ascontiguousarray(self.taxis)).to(device)
        return mask, delay, tax

    @staticmethod
    def merge(slices):
        """Union of slices crossed in one step.  Slices of one bowl have disjoint
        masks, so one mask/delay map carries all of them."""
        first = slices[0]
        if len(slices) == 1:
            return first
        mask = first.mask.copy()
        delays = np.where(first.mask, first.delays, 0.0)
        for s in slices[1:]:
            same = (s.f0 == first.f0 and s.p0 == first.p0 and s.ncycles == first.ncycles
                    and s.dur == first.dur and np.array_equal(s.taxis, first.taxis))
            if not same or np.any(mask & s.mask):
                raise NotImplementedError('source slices crossed together must share pulse '
                                          'parameters and have disjoint masks')
            delays = np.where(s.mask, s.delays, delays)
            mask |= s.mask
        return SourceSlice(mask, delays, first.taxis, first.f0, first.p0, first.ncycles, first.dur)


def bowl_geometry(xaxis, yaxis, c0, radius, roc, dZ=None, inner_radius=0.0, focus=None,
                  element_delays=None, n_elements=1):
    """Aperture, bowl depth map, inverted focusing delays and slice edges (ASS:1126-1166;
    the same geometry and delays as ASS:1023-1056).  ``z_edges`` is None when ``dZ`` is."""
    if focus is None:
        focus = roc
    R = roc
    XX, YY = np.meshgrid(np.asarray(xaxis, dtype=np.float64), np.asarray(yaxis, dtype=np.float64),
                         indexing='ij')
    r2 = XX ** 2 + YY ** 2
    r = np.sqrt(r2)
    aperture = (r2 <= radius ** 2) & (r2 >= inner_radius ** 2)
    bowl_z = R - np.sqrt(R ** 2 - np.minimum(r2, R ** 2 - 1e-20))
    bowl_z[~aperture] = 0.0
    bowl_depth = float(np.max(bowl_z[aperture]))
    dist_to_focus = np.sqrt(r2 + (focus - bowl_z) ** 2)
    delays_full = np.zeros_like(r)
    delays_full[aperture] = dist_to_focus[aperture] / c0
    if n_elements > 1:
        dr = (radius - inner_radius) / n_elements
        for ie in range(n_elements):
            m = (r >= inner_radius + ie * dr) & (r < inner_radius + (ie + 1) * dr) & aperture
            if m.any():
                delays_full[m] = delays_full[m].mean()
                if element_delays is not None:
                    delays_full[m] += np.asarray(element_delays)[ie]
    delays_full[aperture] = delays_full[aperture].max() - delays_full[aperture]
    if dZ is None:
        return aperture, bowl_z, delays_full, None, bowl_depth
    n_slices = max(1, int(np.ceil(bowl_depth / dZ)))
    z_edges = np.linspace(0, bowl_depth, n_slices + 1)
    return aperture, bowl_z, delays_full, z_edges, bowl_depth


def make_bowl_source(xaxis, yaxis, taxis, f0, c0, p0, radius, roc, ncycles=3, dur=2, inner_radius=0.0,
                     focus=None, element_delays=None, n_elements=1, sparse=False, verbose=False):
    """Bowl source on one plane (ASS:971-1077): every aperture pixel fires the tone
    burst at its inverted focusing delay (outermost first).  Returns the dense
    float32 (nX, nY, nT) field, or with ``sparse=True`` its ``SourceSlice``
    (aperture mask + delay map), which the solver synthesises on the GPU."""
    aperture, _, delays, _, _ = bowl_geometry(xaxis, yaxis, c0, radius, roc, None, inner_radius, focus,
                                              element_delays, n_elements)
    delays = np.where(aperture, delays, 0.0)
    sl = SourceSlice(aperture, delays, taxis, f0, p0, ncycles, dur)
    if verbose:
        act = delays[aperture]
        print(f'  Bowl source: outer_r={radius*1e3:.1f} mm, inner_r={inner_radius*1e3:.1f} mm, '
              f'ROC={roc*1e3:.1f} mm, focus={(roc if focus is None else focus)*1e3:.1f} mm, '
              f'{n_elements} elements, {int(aperture.sum())} active grid points, '
              f'delay spread: {(act.max() - act.min())*1e6:.2f} us')
    return sl if sparse else sl.dense()


def make_bowl_source_planes(xaxis, yaxis, taxis, f0, c0, p0, radius, roc, dZ, ncycles=3, dur=2,
                            inner_radius=0.0, focus=None, element_delays=None, n_elements=1,
                            verbose=False):
    """Plane-by-plane bowl source, ASS:1080-1211; returns (planes, bowl_depth)."""
    aperture, bowl_z, delays_full, z_edges, bowl_depth = bowl_geometry(
        xaxis, yaxis, c0, radius, roc, dZ, inner_radius, focus, element_delays, n_elements)
    n_slices = len(z_edges) - 1
    planes = []
    total = 0
    for i in range(n_slices):
        lo, hi = z_edges[i], z_edges[i + 1]
        if i == n_slices - 1:
            m = aperture & (bowl_z >= lo) & (bowl_z <= hi)
        else:
            m = aperture & (bowl_z >= lo) & (bowl_z < hi)
        n_px = int(np.sum(m))
        if n_px == 0:
            continue
        d = delays_full.copy()
        d[m] -= bowl_z[m] / c0
        d[~m] = 0.0
        planes.append((0.5 * (lo + hi), SourceSlice(m, d, taxis, f0, p0, ncycles, dur)))
        total += n_px
    if verbose:
        print(f'  Bowl source (plane-by-plane): bowl_depth={bowl_depth*1e3:.2f} mm, '
              f'{n_slices} z-slices, {total} active grid points')
    return planes, bowl_depth


def generate_phase_screen(nX, nY, dX, dY, c0, f0, correlation_length, speed_std, thickness=None,
                          seed=None):
    """Random sound-speed layer -> phase map (ASS:290-349); returns (phase, c_map) float32."""
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


def skull_screen(nX, nY, dX, dY, f0, seed=3, corr_len=5e-3, t_min=3e-3, t_max=9e-3,
                 c_bone=2800.0, rho_bone=1900.0, c0=1540.0, rho0=1000.0, alpha_db_cm_mhz=8.0):
    """Synthetic skull phase-and-amplitude screen from a seeded Gaussian random field.

    Thickness T = t_min + (t_max - t_min) (g - min g)/(max g - min g) with g of
    spectrum exp(-k^2 L^2/4) (the recipe of ASS:331-337); phase
    omega T (1/c_bone - 1/c0) (form of ASS:345-347); amplitude
    exp(-alpha T) |2 Z_b/(Z_b + Z_0)| (recipe of validate_transcranial.py:203-218).
    """
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nX, dX) * 2 * np.pi
    ky = np.fft.fftfreq(nY, dY) * 2 * np.pi
    K2 = kx[:, None] ** 2 + ky[None, :] ** 2
    noise = rng.standard_normal((nX, nY)) + 1j * rng.standard_normal((nX, nY))
    g = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(np.exp(-K2 * corr_len ** 2 / 4.0))))
    T = t_min + (t_max - t_min) * (g - g.min()) / (g.max() - g.min())
    omega = 2 * np.pi * f0
    phase = omega * T * (1.0 / c_bone - 1.0 / c0)
    alpha_np = alpha_db_cm_mhz * (f0 / 1e6) * 100.0 / (20 * math.log10(math.e))
    zb, z0 = rho_bone * c_bone, rho0 * c0
    amp = np.exp(-alpha_np * T) * abs(2 * zb / (zb + z0))
    return phase.astype(np.float32), amp.astype(np.float32), T
