"""CPU pins of the input builders: the one-plane bowl source (ASS:971-1077), the
skull-CT screens, the sparse-array source and the NRRD reader of the transcranial
workflow (validate_transcranial.py:43-317).

Each builder is checked against closed forms that a dropped term, a wrong sign or
a transposed axis would break (bowl-surface geometry, interface transmission,
exactness of trilinear sampling on a linear volume) and against the oracle's
restatement where the reference's own loop structure is the thing to match
(per-pixel element loop, overwrite order at overlaps).
"""
import gzip
import os

import numpy as np
import pytest

from angularspectrum_heterogeneous_amd import (SourceSlice, build_initial_condition, load_skull_slab,
                                               make_bowl_source, make_bowl_source_planes, read_nrrd,
                                               screens_to_xz_maps, skull_to_phase_screens)
from angularspectrum_heterogeneous_amd.sources import bowl_geometry
from angularspectrum_heterogeneous_amd.transcranial import hu_to_acoustic
from oracle import asm_oracle as O

C0 = 1540.0


def _axes(n, d):
    return (np.arange(n) - (n - 1) / 2) * d


# ------------------------------------------------------------ bowl source ---
def test_bowl_source_matches_oracle_and_sparse_form():
    x = _axes(40, 5e-4)
    t = (np.arange(96) - 47.5) * 4e-8
    kw = dict(inner_radius=1e-3, focus=14e-3, n_elements=3, element_delays=[0.0, 1e-7, -2e-7])
    a = make_bowl_source(x, x, t, 1e6, C0, 2e5, 8e-3, 20e-3, ncycles=3, dur=2, **kw)
    b = O.make_bowl_source(x, x, t, 1e6, C0, 2e5, 8e-3, 20e-3, ncycles=3, dur=2, **kw)
    assert a.dtype == np.float32 and a.shape == (40, 40, 96)
    np.testing.assert_array_equal(a, b)
    sl = make_bowl_source(x, x, t, 1e6, C0, 2e5, 8e-3, 20e-3, ncycles=3, dur=2, sparse=True, **kw)
    assert isinstance(sl, SourceSlice)
    np.testing.assert_array_equal(np.asarray(sl), a)


def test_bowl_at_geometric_focus_fires_all_pixels_together():
    # every point of a sphere of radius R is R from its centre: zero delay spread
    x = _axes(48, 4e-4)
    t = (np.arange(64) - 31.5) * 4e-8
    sl = make_bowl_source(x, x, t, 1e6, C0, 1e5, 9e-3, 25e-3, sparse=True)
    assert sl.mask.sum() > 100
    assert np.max(np.abs(sl.delays[sl.mask])) < 1e-15
    f = np.asarray(sl)
    cols = f[sl.mask]
    assert np.abs(cols[0]).max() > 0.5e5
    np.testing.assert_allclose(cols, np.broadcast_to(cols[0], cols.shape), rtol=0, atol=1e-6 * 1e5)


def test_bowl_electronic_focus_delays_closed_form():
    # |P - F|^2 = F^2 + 2 z (R - F) for a bowl point P at sagitta z (R the curvature radius)
    x = _axes(50, 4e-4)
    R, F, a = 30e-3, 18e-3, 9e-3
    sl = make_bowl_source(x, x, np.zeros(4), 1e6, C0, 1.0, a, R, focus=F, sparse=True)
    r2 = x[:, None] ** 2 + x[None, :] ** 2
    z = r2 / (R + np.sqrt(R * R - r2))           # R - sqrt(R^2 - r^2), cancellation-free
    d = np.sqrt(F * F + 2 * z * (R - F))
    want = (d[sl.mask].max() - d[sl.mask]) / C0  # farthest (rim) fires first
    np.testing.assert_allclose(sl.delays[sl.mask], want, rtol=0, atol=1e-16)
    rr = np.sqrt(r2)[sl.mask]
    assert sl.delays[sl.mask][np.argmax(rr)] == 0.0
    assert sl.delays[sl.mask][np.argmin(rr)] == pytest.approx(want.max(), rel=1e-12)


def test_bowl_rings_and_steering_delays():
    x = _axes(44, 5e-4)
    base = dict(inner_radius=2e-3, focus=15e-3, n_elements=3)
    ed = np.array([0.0, 3e-7, 1e-7])
    s0 = make_bowl_source(x, x, np.zeros(4), 1e6, C0, 1.0, 9e-3, 20e-3, sparse=True, **base)
    s1 = make_bowl_source(x, x, np.zeros(4), 1e6, C0, 1.0, 9e-3, 20e-3, sparse=True, element_delays=ed, **base)
    r = np.sqrt(x[:, None] ** 2 + x[None, :] ** 2)
    dr = (9e-3 - 2e-3) / 3
    ring = [(r >= 2e-3 + i * dr) & (r < 2e-3 + (i + 1) * dr) & s0.mask for i in range(3)]
    v0 = [np.unique(s0.delays[m]) for m in ring]
    v1 = [np.unique(s1.delays[m]) for m in ring]
    assert all(len(v) == 1 for v in v0 + v1)
    # the inversion max - d turns an added steering delay into an earlier firing
    for i in range(3):
        for j in range(3):
            assert (v1[i][0] - v1[j][0]) == pytest.approx((v0[i][0] - v0[j][0]) - (ed[i] - ed[j]), abs=1e-15)


def test_bowl_planes_agree_with_one_plane_source():
    # slice delay + bowl depth / c0 equals the one-plane delay; slices partition the aperture
    x = _axes(40, 5e-4)
    t = (np.arange(32) - 15.5) * 4e-8
    kw = dict(inner_radius=1e-3, focus=14e-3, n_elements=2)
    planes, _ = make_bowl_source_planes(x, x, t, 1e6, C0, 1.0, 8e-3, 20e-3, 4e-4, **kw)
    flat = make_bowl_source(x, x, t, 1e6, C0, 1.0, 8e-3, 20e-3, sparse=True, **kw)
    ap, bz, _, _, _ = bowl_geometry(x, x, C0, 8e-3, 20e-3, None, **kw)
    cover = np.zeros(ap.shape, int)
    for _, sl in planes:
        cover += sl.mask
        np.testing.assert_allclose(sl.delays[sl.mask] + bz[sl.mask] / C0, flat.delays[sl.mask], rtol=0,
                                   atol=1e-18)
    assert np.array_equal(cover, ap.astype(int))


# ------------------------------------------------------------ skull screens ---
def test_skull_screens_closed_forms():
    f0, dz = 0.5e6, 2e-4
    nx, ny = 6, 5
    c = np.full((nx, ny, 4), C0)
    rho = np.full((nx, ny, 4), 1000.0)
    c[:, :, 1], rho[:, :, 1] = 2900.0, 2200.0       # full bone
    c[:, :, 2], rho[:, :, 2] = 2220.0, 1600.0       # half bone
    c[2, 3, 3] = 1600.0                             # one pixel of a near-water slice
    scr = skull_to_phase_screens(c, dz, 3e-4, C0, f0, rho_map=rho, rho0=1000.0)
    assert [s[0] for s in scr] == [1 * dz, 2 * dz, 3 * dz]   # the water slice is left out
    w = 2 * np.pi * f0
    for (z, ph, amp), cb, rb, frac in [(scr[0], 2900.0, 2200.0, 1.0), (scr[1], 2220.0, 1600.0, 0.5)]:
        assert ph.dtype == np.float32 and amp.dtype == np.float32
        np.testing.assert_allclose(ph, w * dz * (C0 - cb) / (cb * C0), rtol=1e-6)
        alpha_np = (0.5 + frac * 7.5) * (f0 / 1e6) * 11.51          # dB/cm/MHz -> Np/m as rounded there
        tr = 2 * rb * cb / (rb * cb + 1000.0 * C0)                   # normal-incidence transmission
        np.testing.assert_allclose(amp, np.exp(-alpha_np * dz) * tr, rtol=1e-6)
    ph3 = scr[2][1]
    assert ph3[2, 3] == pytest.approx(w * dz * (1 / 1600.0 - 1 / C0), rel=1e-6)
    assert np.count_nonzero(ph3) == 1
    # without a density map the amplitude is the attenuation alone; sub-tissue speeds clip to tissue
    c2 = np.full((2, 2, 1), 1400.0)
    (z, ph, amp), = skull_to_phase_screens(c2, dz, dz, C0, f0)
    np.testing.assert_allclose(amp, np.exp(-0.5 * 0.5 * 11.51 * dz), rtol=1e-6)
    np.testing.assert_allclose(ph, w * dz * (C0 - 1400.0) / (1400.0 * C0), rtol=1e-6)


def test_screens_to_xz_maps():
    assert screens_to_xz_maps([], 1) == (None, None)
    ph = np.arange(12, dtype=np.float32).reshape(4, 3)
    amp = np.full((4, 3), 0.1, np.float32)
    amp[:, 1] = 0.0
    p, a = screens_to_xz_maps([(0.0, ph, amp), (1.0, 2 * ph, amp)], 1)
    np.testing.assert_array_equal(p, np.stack([ph[:, 1], 2 * ph[:, 1]]))
    np.testing.assert_allclose(a, 600.0, rtol=1e-6)   # amplitude floor 1e-30 -> 600 dB (float32)


# ------------------------------------------------------------ sparse array ---
def test_sparse_array_matches_per_pixel_loop():
    rng = np.random.default_rng(7)
    n, d = 36, 2.5e-4                       # 3x3 pixel elements
    x = _axes(n, d)
    t = (np.arange(48) - 23.5) * 1e-7
    pos = np.column_stack([rng.uniform(-5e-3, 5e-3, 30), rng.uniform(-5e-3, 5e-3, 30), np.zeros(30)])
    pos[5] = pos[4] + np.array([d, 0, 0])   # overlapping squares: the later element wins
    pos[6] = np.array([-4.4e-3, 4.4e-3, 0])  # clipped at the grid edge
    ids = np.arange(28)                      # only the first len(ids) positions are used
    a = build_initial_condition(pos, ids, x, x, t, 0.5e6, C0, 1e5, 30e-3)
    b = O.build_initial_condition(pos, ids, x, x, t, 0.5e6, C0, 1e5, 30e-3)
    np.testing.assert_array_equal(a, b)
    sl = build_initial_condition(pos, ids, x, x, t, 0.5e6, C0, 1e5, 30e-3, sparse=True)
    np.testing.assert_array_equal(np.asarray(sl), a)


def test_sparse_array_focusing_delays_closed_form():
    n, d = 41, 2.5e-4
    x = _axes(n, d)
    k = 6
    # four elements on a circle of 6 pixels around the centre and one at the centre
    pos = np.array([[0, 0], [k * d, 0], [-k * d, 0], [0, k * d], [0, -k * d]], float)
    pos = np.column_stack([pos, np.zeros(5)])
    F = 20e-3
    sl = build_initial_condition(pos, np.arange(5), x, x, np.zeros(8), 0.5e6, C0, 1.0, F, sparse=True)
    c = n // 2
    ring = [(c + k, c), (c - k, c), (c, c + k), (c, c - k)]
    for i, j in ring:
        assert sl.delays[i, j] == 0.0 and sl.mask[i - 1:i + 2, j - 1:j + 2].all()
    want = (np.sqrt(F * F + (k * d) ** 2) - F) / C0
    assert sl.delays[c, c] == pytest.approx(np.float32(want), rel=1e-7)
    assert sl.mask.sum() == 5 * 9


# -------------------------------------------------------------------- NRRD ---
def _write_nrrd(path, data, voxel_mm, encoding="raw", endian="little", detached=False):
    arr = np.asarray(data)
    kind = {np.dtype(np.int16): "int16", np.dtype(np.float32): "float", np.dtype(np.uint8): "uint8"}[arr.dtype]
    raw = arr.astype(arr.dtype.newbyteorder("<" if endian == "little" else ">")).tobytes(order="F")
    if encoding == "gzip":
        raw = gzip.compress(raw)
    dirs = " ".join("(" + ",".join(str(voxel_mm if i == j else 0.0) for j in range(3)) + ")" for i in range(3))
    head = ["NRRD0004", "# test volume", f"type: {kind}", "dimension: 3", "space: left-posterior-superior",
            "sizes: " + " ".join(str(s) for s in arr.shape), f"space directions: {dirs}",
            "kinds: domain domain domain", f"endian: {endian}", f"encoding: {encoding}",
            "space origin: (0,0,0)", "comment:=key value pairs are skipped"]
    if detached:
        head.append("data file: " + os.path.basename(path) + ".raw")
        with open(path + ".raw", "wb") as g:
            g.write(raw)
        raw = b""
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n\n").encode())
        f.write(raw)


@pytest.mark.parametrize("encoding,endian,detached", [("raw", "little", False), ("gzip", "little", False),
                                                      ("raw", "big", False), ("raw", "little", True)])
def test_nrrd_reader_round_trip(tmp_path, encoding, endian, detached):
    rng = np.random.default_rng(1)
    vol = rng.integers(-1000, 3000, size=(7, 5, 3)).astype(np.int16)
    p = str(tmp_path / "v.nrrd")
    _write_nrrd(p, vol, 0.25, encoding, endian, detached)
    data, hdr = read_nrrd(p)
    assert data.shape == (7, 5, 3) and hdr["sizes"] == [7, 5, 3]
    np.testing.assert_array_equal(data, vol)
    np.testing.assert_array_equal(hdr["space directions"], 0.25 * np.eye(3))
    # axis 0 is the fastest in the file: the first bytes are vol[0, 0, 0], vol[1, 0, 0]
    if not detached and encoding == "raw":
        body = open(p, "rb").read().split(b"\n\n", 1)[1]
        first = np.frombuffer(body[:4], dtype=("<" if endian == "little" else ">") + "i2")
        np.testing.assert_array_equal(first, vol[:2, 0, 0])


def test_nrrd_reader_rejects_what_it_does_not_read(tmp_path):
    p = tmp_path / "bad.nrrd"
    p.write_bytes(b"NRRD0004\ntype: short\ndimension: 1\nsizes: 4\nencoding: bzip2\n\n\x00" * 1)
    with pytest.raises(ValueError, match="encoding"):
        read_nrrd(str(p))
    q = tmp_path / "notnrrd.nrrd"
    q.write_bytes(b"PK\x03\x04")
    with pytest.raises(ValueError, match="not an NRRD"):
        read_nrrd(str(q))


def test_skull_slab_samples_a_linear_volume_exactly(tmp_path):
    # trilinear interpolation is exact on a linear function of (L, P, S)
    vox = 0.5
    n = (60, 60, 40)
    L, P, S = np.meshgrid(*[np.arange(k) * vox for k in n], indexing="ij")
    hu = (400.0 + 80.0 * L - 25.0 * P + 30.0 * S).astype(np.float32)
    path = str(tmp_path / "ct.nrrd")
    _write_nrrd(path, hu, vox)
    place = {"xdc_center_lps": (15.0, 14.0, 10.0), "normal": (0.6, 0.8), "tangent": (0.8, -0.6), "z_mm": 10.0}
    size = (8e-3, 6e-3, 20e-3)   # the far end of the slab leaves the volume (0 HU there)
    c, rho, dz = load_skull_slab(path, place, size, 0.5e-3)
    assert c.shape == (16, 12, 40) and c.dtype == np.float32 and rho.dtype == np.float32
    assert dz == pytest.approx(20e-3 / 39)
    lat = np.linspace(-4, 4, 16)
    elev = np.linspace(0, 6, 12) - 3.0
    dep = np.linspace(0, 20, 40)
    Lp = 15.0 + lat[:, None, None] * 0.8 + dep[None, None, :] * 0.6 + 0 * elev[None, :, None]
    Pp = 14.0 - lat[:, None, None] * 0.6 + dep[None, None, :] * 0.8 + 0 * elev[None, :, None]
    Sp = 10.0 + elev[None, :, None] + 0 * lat[:, None, None] + 0 * dep[None, None, :]
    inside = ((Lp >= 0) & (Lp <= (n[0] - 1) * vox) & (Pp >= 0) & (Pp <= (n[1] - 1) * vox)
              & (Sp >= 0) & (Sp <= (n[2] - 1) * vox))
    want = np.where(inside, 400.0 + 80.0 * Lp - 25.0 * Pp + 30.0 * Sp, 0.0)
    cw, rw = hu_to_acoustic(want)
    assert inside.any() and (~inside).any() and (cw == 2900).any() and ((cw > 1540) & (cw < 2900)).any()
    np.testing.assert_allclose(c, cw, rtol=2e-6)
    np.testing.assert_allclose(rho, rw, rtol=2e-6)
    # mapping end points: tissue below 700 HU, bone above 1973 HU
    cc, rr = hu_to_acoustic(np.array([0.0, 700.0, 1336.5, 1973.0, 3000.0]))
    np.testing.assert_allclose(cc, [1540, 1540, 2220, 2900, 2900])
    np.testing.assert_allclose(rr, [1000, 1000, 1600, 2200, 2200])
