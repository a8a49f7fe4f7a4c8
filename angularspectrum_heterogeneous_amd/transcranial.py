"""Transcranial inputs: skull CT -> phase/amplitude screens and the sparse-array source.

Host-side builders of the inputs the march consumes, restated from the
reference's transcranial workflow (validate_transcranial.py):

* ``read_nrrd`` reads an NRRD volume (the microCT format the workflow loads
  with pynrrd, which is not a dependency here): attached or detached data,
  raw or gzip encoding, returning the array in the file's axis order
  (fastest axis first, as pynrrd's default ``index_order='F'`` does) and a
  header dict whose ``'space directions'`` is a float array;
* ``load_skull_slab`` samples a beam-aligned slab of the CT and maps
  Hounsfield units to sound speed and density (validate_transcranial.py:43-142);
* ``skull_to_phase_screens`` turns (c, rho) maps into the
  ``SolverParams.phaseScreens`` list (validate_transcranial.py:166-238);
* ``screens_to_xz_maps`` (validate_transcranial.py:241-250);
* ``load_sparse_connector`` reads the array layout from a MATLAB v7.3 file
  (validate_transcranial.py:253-258); it needs h5py and raises ImportError
  without it;
* ``build_initial_condition`` maps element positions onto the grid with
  geometric focusing delays (validate_transcranial.py:261-317); with
  ``sparse=True`` it returns the ``SourceSlice`` (mask + delay map) that the
  solver synthesises on the GPU instead of a dense (nX, nY, nT) array.
"""
from __future__ import annotations

import gzip
import os
import re
from typing import List, Optional, Tuple

import numpy as np

from .sources import SourceSlice

# --------------------------------------------------------------------------
# NRRD
# --------------------------------------------------------------------------
_NRRD_TYPES = {
    "int8": "i1", "signed char": "i1", "int8_t": "i1",
    "uint8": "u1", "uchar": "u1", "unsigned char": "u1", "uint8_t": "u1",
    "int16": "i2", "short": "i2", "short int": "i2", "signed short": "i2", "signed short int": "i2",
    "int16_t": "i2",
    "uint16": "u2", "ushort": "u2", "unsigned short": "u2", "unsigned short int": "u2", "uint16_t": "u2",
    "int32": "i4", "int": "i4", "signed int": "i4", "int32_t": "i4",
    "uint32": "u4", "uint": "u4", "unsigned int": "u4", "uint32_t": "u4",
    "int64": "i8", "longlong": "i8", "long long": "i8", "long long int": "i8", "signed long long": "i8",
    "signed long long int": "i8", "int64_t": "i8",
    "uint64": "u8", "ulonglong": "u8", "unsigned long long": "u8", "unsigned long long int": "u8",
    "uint64_t": "u8",
    "float": "f4", "double": "f8",
}


def _parse_vector(tok: str):
    tok = tok.strip()
    if tok == "none":
        return None
    if not (tok.startswith("(") and tok.endswith(")")):
        raise ValueError(f"NRRD: bad vector {tok!r}")
    return [float(v) for v in tok[1:-1].split(",")]


def read_nrrd(path: str) -> Tuple[np.ndarray, dict]:
    """Reads an NRRD file; returns (data, header) with data.shape == header['sizes']."""
    with open(path, "rb") as f:
        magic = f.readline().decode("ascii", "replace").strip()
        if not magic.startswith("NRRD000"):
            raise ValueError(f"{path}: not an NRRD file")
        header: dict = {}
        while True:
            line = f.readline()
            if not line:
                break
            text = line.decode("latin-1").rstrip("\r\n")
            if text == "":
                break
            if text.startswith("#") or ":=" in text:
                continue
            key, _, val = text.partition(": ")
            header[key.strip()] = val.strip()
        body = f.read()
    if "type" not in header or "sizes" not in header or "encoding" not in header:
        raise ValueError(f"{path}: NRRD header lacks type, sizes or encoding")
    dt = _NRRD_TYPES.get(header["type"].lower())
    if dt is None:
        raise ValueError(f"{path}: unsupported NRRD type {header['type']!r}")
    sizes = [int(v) for v in header["sizes"].split()]
    endian = header.get("endian", "little")
    dtype = np.dtype(("<" if endian == "little" else ">") + dt)
    data_file = header.get("data file", header.get("datafile"))
    if data_file is not None:
        if data_file.startswith("LIST") or "%" in data_file:
            raise ValueError(f"{path}: multi-file NRRD data is not supported")
        with open(os.path.join(os.path.dirname(os.path.abspath(path)), data_file), "rb") as g:
            body = g.read()
    enc = header["encoding"].lower()
    if enc in ("gzip", "gz"):
        body = gzip.decompress(body)
    elif enc != "raw":
        raise ValueError(f"{path}: unsupported NRRD encoding {enc!r}")
    n = int(np.prod(sizes))
    skip = int(header.get("byte skip", "0"))
    if skip == -1:
        skip = len(body) - n * dtype.itemsize
    raw = np.frombuffer(body, dtype=dtype, count=n, offset=skip)
    data = raw.reshape(sizes, order="F")
    header["sizes"] = sizes
    if "space directions" in header:
        vecs = [_parse_vector(t) for t in re.findall(r"\([^)]*\)|none", header["space directions"])]
        width = max((len(v) for v in vecs if v is not None), default=0)
        header["space directions"] = np.array(
            [v if v is not None else [np.nan] * width for v in vecs], dtype=np.float64)
    if "space origin" in header:
        header["space origin"] = np.array(_parse_vector(header["space origin"]), dtype=np.float64)
    if "spacings" in header:
        header["spacings"] = np.array([float(v) for v in header["spacings"].split()], dtype=np.float64)
    return data, header


# --------------------------------------------------------------------------
# Skull CT -> acoustic maps -> screens
# --------------------------------------------------------------------------
HU_LOWER, HU_UPPER = 700.0, 1973.0
C_TISSUE, C_BONE = 1540.0, 2900.0
RHO_TISSUE, RHO_BONE = 1000.0, 2200.0


def hu_to_acoustic(hu: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Hounsfield units -> (c, rho) float32, linear in the bone fraction between
    700 and 1973 HU (validate_transcranial.py:126-133)."""
    frac = np.clip((hu - HU_LOWER) / (HU_UPPER - HU_LOWER), 0, 1)
    c = (C_TISSUE + frac * (C_BONE - C_TISSUE)).astype(np.float32)
    rho = (RHO_TISSUE + frac * (RHO_BONE - RHO_TISSUE)).astype(np.float32)
    return c, rho


def load_skull_slab(nrrd_path: str, placement: dict, slab_size_m, dx_target_m: float):
    """Beam-aligned skull slab from a microCT (validate_transcranial.py:43-142).

    ``placement`` holds ``xdc_center_lps`` (mm), ``normal`` and ``tangent``
    (in-plane L, P unit vectors) and ``z_mm``; ``slab_size_m`` is
    (lateral, elevational, depth).  The CT is sampled trilinearly on the slab
    grid (lateral centred, elevation and depth from 0), points outside the
    volume count as 0 HU.  Returns (c_map, rho_map, dz_slab).
    """
    from scipy.interpolate import RegularGridInterpolator

    data, header = read_nrrd(nrrd_path)
    voxel_mm = float(header["space directions"][0, 0])
    axes = tuple(np.arange(data.shape[d]) * voxel_mm for d in range(3))
    interp = RegularGridInterpolator(axes, data.astype(np.float64), method="linear", bounds_error=False,
                                     fill_value=np.nan)
    xdc_l, xdc_p, _ = placement["xdc_center_lps"]
    norm_l, norm_p = placement["normal"]
    tang_l, tang_p = placement["tangent"]
    z_center_mm = placement["z_mm"]
    w_lat, w_elev, w_depth = slab_size_m
    n_lat = int(np.round(w_lat / dx_target_m))
    n_elev = int(np.round(w_elev / dx_target_m))
    n_depth = int(np.round(w_depth / dx_target_m))
    lat_mm = np.linspace(-w_lat / 2, w_lat / 2, n_lat) * 1e3
    elev_mm = np.linspace(0, w_elev, n_elev) * 1e3
    depth = np.linspace(0, w_depth, n_depth)
    lat_g, dep_g = np.meshgrid(lat_mm, depth * 1e3, indexing="ij")
    l_pts = xdc_l + lat_g * tang_l + dep_g * norm_l
    p_pts = xdc_p + lat_g * tang_p + dep_g * norm_p
    hu = np.empty((n_lat, n_elev, n_depth), dtype=np.float64)
    for j in range(n_elev):
        s_pts = np.full_like(l_pts, z_center_mm + (elev_mm[j] - w_elev * 1e3 / 2))
        pts = np.stack([l_pts.ravel(), p_pts.ravel(), s_pts.ravel()], axis=-1)
        hu[:, j, :] = interp(pts).reshape(n_lat, n_depth)
    hu = np.nan_to_num(hu, nan=0.0)
    c_map, rho_map = hu_to_acoustic(hu)
    return c_map, rho_map, depth[1] - depth[0]


def skull_to_phase_screens(c_map, dz_screen, dz_prop, c0, f0, rho_map=None, rho0=1000.0, alpha_bone=8.0,
                           alpha_tissue=0.5, verbose=False) -> List[tuple]:
    """(nx, ny, nz) sound speed (and density) -> [(z, phase, amplitude), ...]
    (validate_transcranial.py:166-238).

    Slice iz is a screen at z = iz * dz_screen with phase
    omega dz_screen (1/c - 1/c0) and amplitude exp(-alpha dz_screen) (times
    |2 Z/(Z + Z0)| with a density map), alpha interpolated between tissue and
    bone by (c - 1540)/(2900 - 1540) in dB/cm/MHz, 1 dB/cm = 11.51 Np/m as the
    workflow rounds it; slices with max|phase| <= 1e-6 and min amplitude
    >= 0.999 are left out.  ``dz_prop`` is accepted for the workflow's
    signature and not used, as there.
    """
    c_map = np.asarray(c_map)
    nx, ny, nz = c_map.shape
    omega = 2 * np.pi * f0
    f_mhz = f0 / 1e6
    screens = []
    for iz in range(nz):
        c = c_map[:, :, iz]
        phase = omega * dz_screen * (1.0 / c - 1.0 / c0)
        bone = np.clip((c - 1540.0) / (2900.0 - 1540.0), 0, 1)
        alpha = alpha_tissue + bone * (alpha_bone - alpha_tissue)
        amp = np.exp(-(alpha * f_mhz * 0.1151 / 0.01) * dz_screen)
        if rho_map is not None:
            z_loc = np.asarray(rho_map)[:, :, iz] * c
            amp = amp * np.abs(2.0 * z_loc / (z_loc + rho0 * c0))
        if np.max(np.abs(phase)) > 1e-6 or np.min(amp) < 0.999:
            screens.append((iz * dz_screen, phase.astype(np.float32), amp.astype(np.float32)))
    if verbose:
        print(f'  Generated {len(screens)} screens from {nz} z-slices')
    return screens


def screens_to_xz_maps(screens, mid_elev: int):
    """Phase (rad) and attenuation (dB) x-z maps at one elevation row
    (validate_transcranial.py:241-250); (None, None) without screens."""
    if not screens:
        return None, None
    phase_xz = np.stack([ph[:, mid_elev] for _, ph, _ in screens], axis=0)
    atten_db = np.stack([-20.0 * np.log10(np.maximum(a[:, mid_elev], 1e-30)) for _, _, a in screens], axis=0)
    return phase_xz, atten_db


# --------------------------------------------------------------------------
# Sparse array source
# --------------------------------------------------------------------------
def load_sparse_connector(mat_path: str):
    """Element positions (n, 3) and element ids from a MATLAB v7.3 (HDF5) file
    (validate_transcranial.py:253-258); needs h5py."""
    try:
        import h5py
    except ImportError as e:
        raise ImportError('load_sparse_connector reads MATLAB v7.3 (HDF5) files and needs h5py') from e
    with h5py.File(mat_path, 'r') as f:
        positions = f['Position'][()].T
        element_ids = f['SparseConnector'][()].ravel().astype(int)
    return positions, element_ids


def build_initial_condition(positions, element_ids, xaxis, yaxis, t, f0, c0, p0, focal_depth, ncycles=2,
                            sparse=False, verbose=False):
    """Focused sparse-array source field (validate_transcranial.py:261-317).

    Element m (m < len(element_ids)) sits at grid index
    round(pos/dX) - int(mean(round(pos/dX))) + n//2 along x and y and covers a
    square of max(1, round(750e-6/dX)) pixels (clipped at the grid edge; a
    later element overwrites an earlier one where squares overlap).  Its delay
    is (max dist - dist)/c0, dist from the element to the point focal_depth
    below the grid centre; the delay map is stored in float32 as there.
    Returns the dense float32 (nX, nY, nT) field, or its ``SourceSlice``
    with ``sparse=True``.
    """
    positions = np.asarray(positions)
    nX, nY = len(xaxis), len(yaxis)
    dX = xaxis[1] - xaxis[0]
    iy_raw = np.round(positions[:, 0] / dX).astype(int)
    iz_raw = np.round(positions[:, 1] / dX).astype(int)
    iy_c = iy_raw - int(np.mean(iy_raw)) + nX // 2
    iz_c = iz_raw - int(np.mean(iz_raw)) + nY // 2
    fy, fz = nX // 2, nY // 2
    dist = np.sqrt(focal_depth ** 2 + (iy_c * dX - fy * dX) ** 2 + (iz_c * dX - fz * dX) ** 2)
    delays = (dist.max() - dist) / c0
    half = max(1, round(750e-6 / dX)) // 2
    ap = np.zeros((nX, nY), dtype=bool)
    tt = np.zeros((nX, nY), dtype=np.float32)
    for m in range(len(element_ids)):
        y0, y1 = max(iy_c[m] - half, 0), min(iy_c[m] + half + 1, nX)
        z0, z1 = max(iz_c[m] - half, 0), min(iz_c[m] + half + 1, nY)
        if y0 < y1 and z0 < z1:
            ap[y0:y1, z0:z1] = True
            tt[y0:y1, z0:z1] = delays[m]
    sl = SourceSlice(ap, tt.astype(np.float64), t, f0, p0, ncycles, 2)
    if verbose:
        print(f'  Sparse array: {int(ap.sum())} active grid points, '
              f'delay spread: {(delays.max() - delays.min()) * 1e6:.1f} us')
    return sl if sparse else sl.dense()


__all__ = ["read_nrrd", "hu_to_acoustic", "load_skull_slab", "skull_to_phase_screens", "screens_to_xz_maps",
           "load_sparse_connector", "build_initial_condition"]
