"""Restart onto another grid on the device (tc_regrid_field: ReadFlowField
with interp.F90) against the oracle's restatement, for each field's
staggering, refining up to 2x and coarsening below 2x, with ISTR 0 and 6
radial grids.  Bar: 1e-13 of the largest old value (the two sides differ
only by FMA contraction)."""
import numpy as np
import pytest

from helpers import small_params

pytestmark = pytest.mark.gpu

CASES = [
    # (new nthm, nzm, nrm, istr, str), (old nthm, nzm, nr, istro, stro)
    ((16, 12, 16, 6, 2.0), (8, 6, 9, 0, 0.0)),
    ((10, 8, 12, 0, 0.0), (12, 10, 13, 6, 1.0)),
    ((48, 40, 24, 6, 3.0), (32, 24, 17, 6, 3.0)),
]


@pytest.mark.parametrize("case", CASES)
def test_regrid_matches_oracle(tc, oracle, case):
    (nthm, nzm, nrm, istr, st), (tho, zo, ro, istro, stro) = case
    p = small_params(tc, nthm, nzm, nrm, istr=istr, str=st)
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    rng = np.random.default_rng(nthm + nzm)
    for name in ("qt", "qr", "qz"):
        old = rng.uniform(-1, 1, (ro, zo, tho))
        dev.regrid(name, old, istro, stro)
        want = orc.regrid(name, old, istro, stro)
        got = dev.get(name)
        err = np.abs(got - want).max()
        assert err <= 1e-13 * np.abs(old).max(), (name, err)


def test_regrid_rejects(tc):
    p = small_params(tc, 16, 12, 16, istr=6, str=2.0)
    dev = tc.Solver(p)
    with pytest.raises(tc.TCError):
        dev.regrid("qt", np.zeros((9, 5, 7)), 0, 0.0)    # new nth 17 > 2 * 8
    with pytest.raises(tc.TCError):
        dev.regrid("pr", np.zeros((9, 11, 15)), 0, 0.0)  # pressure is not read back
    with pytest.raises(tc.TCError):
        dev.regrid("qt", np.zeros((9, 11, 15)), 4, 1.5)  # gridnew builds no ISTR 4 grid
