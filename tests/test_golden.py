"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py).

The FP64 oracle's outputs of two small seeded marches are committed as
regression anchors: the oracle must keep reproducing them (CPU), and the GPU
march must match them per output within rel-L2 1e-4 (the single-precision march
against FP64), with ``zaxis`` equal.  The cases cross a phase/amplitude screen
and a dense source plane, run the frequency-weighted and super-absorbing
boundaries, the sequential water branch with TVD Rusanov, and end on a partial step.
"""
import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.golden import make_fixtures as G

NAMES = ("pnp", "ppp", "pI", "pIloss", "pax", "field_xc", "field_yc")


@pytest.mark.parametrize("case", G.CASES)
def test_oracle_reproduces_golden_fixture(case):
    want = np.load(G.path(case))
    got = G.oracle_outputs(case)
    np.testing.assert_array_equal(got["zaxis"], want["zaxis"])
    for n in NAMES:
        assert got[n].shape == want[n].shape, n
        assert rel_l2(got[n], want[n]) < 1e-6, n  # float32 storage of FP64 values


@pytest.mark.gpu
@pytest.mark.parametrize("case", G.CASES)
def test_gpu_march_matches_golden_fixture(case):
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    print(f"[golden] {case}", flush=True)
    want = np.load(G.path(case))
    field, kw = G.case_inputs(case)
    got = G.summarize(angular_spectrum_solve(field, SolverParams(**kw), verbose=False))
    np.testing.assert_allclose(got["zaxis"], want["zaxis"], rtol=0, atol=1e-15)
    errs = {n: rel_l2(got[n], want[n]) for n in NAMES}
    print(errs, flush=True)
    for n, e in errs.items():
        assert e <= 1e-4, (n, e)
