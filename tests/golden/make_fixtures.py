"""Writes the golden fixtures of this directory: FP64-oracle outputs of small seeded marches.

The reference ships no golden vectors (its validation scripts write git-ignored
``.npz`` files), so these fixtures come from the project's FP64 oracle
(``oracle/asm_oracle.py``): regression anchors for the oracle itself (the CPU
test regenerates and compares them) and FP64 expectations for the GPU march
that need no oracle run at test time.  Each case is rebuilt from code and a
seed by ``case_inputs``; the fixture stores the per-step outputs, ``zaxis``,
``pax`` and the final field's two centre planes.

    python tests/golden/make_fixtures.py      # rewrites tests/golden/*.npz
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

CASES = ("gauss_fw_sa_screen_source", "sequential_water")


def case_inputs(name: str):
    """(initial field float32 (64, 64, 256), SolverParams keyword arguments)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from angularspectrum_heterogeneous_amd import workloads as W
    field, kw = W.config1()
    nX, nY, nT = field.shape
    rng = np.random.default_rng(2024)
    if name == "gauss_fw_sa_screen_source":
        phase = rng.uniform(-1.5, 1.5, (nX, nY)).astype(np.float32)
        amp = rng.uniform(0.4, 1.1, (nX, nY)).astype(np.float32)
        src = (0.25 * W.gaussian_beam(p0=1e5, sigma=3e-3)).astype(np.float32)
        kw.update(propDist=0.0029, useFreqWeightedBoundary=True, useSuperAbsorbing=True,
                  phaseScreens=[(1.1e-3, phase, amp)], sourcePlanes=[(1.9e-3, src)])
    elif name == "sequential_water":
        kw.update(useSplitStep=False, alpha0=-1.0, propDist=0.0025, fluxScheme='rusanov', useTVD=True,
                  boundaryProfile='quadratic')
    else:
        raise KeyError(name)
    return field, kw


def summarize(out) -> dict:
    field, pnp, ppp, pI, pIloss, zaxis, pax = out
    f = np.asarray(field)
    c = f.shape[0] // 2
    return {"pnp": np.asarray(pnp, np.float32), "ppp": np.asarray(ppp, np.float32),
            "pI": np.asarray(pI, np.float32), "pIloss": np.asarray(pIloss, np.float32),
            "zaxis": np.asarray(zaxis, np.float64), "pax": np.asarray(pax, np.float32),
            "field_xc": f[c].astype(np.float32), "field_yc": f[:, c].astype(np.float32)}


def oracle_outputs(name: str) -> dict:
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import asm_oracle as O
    field, kw = case_inputs(name)
    return summarize(O.angular_spectrum_solve(field, O.OracleParams(**kw), dtype=np.float64))


def path(name: str) -> str:
    return os.path.join(HERE, name + ".npz")


def main() -> None:
    for name in CASES:
        np.savez_compressed(path(name), **oracle_outputs(name))
        print(name, os.path.getsize(path(name)), "bytes")


if __name__ == "__main__":
    main()
