"""Grids that do not fit in HBM: context creation and batch admission fail
with an error (the allocation that failed is named), release what they had
allocated, and leave the device usable for the next context or batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# 2048 x 4096 x 8193: one field alone is 550 GB, beyond one MI355X's
# 288 GB (FFT lengths the plans accept, so creation reaches the allocations)
HUGE = (2048, 4096, 8193)


def _params(tc, nthm, nzm, nrm):
    p = tc.configs.scaled(tc.configs.small_64(), nthm, nzm, nrm)
    p.update(idtv=0, dtmax=2e-3, dtmin=2e-3, limitcfl=1e6, tout=1e30, tmax=1e30, ntst=100)
    return p


def _steps_ok(tc, s, p):
    qt, qr, qz = tc.init_perturbed_couette(p, 3, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    s.project()
    for n in range(2):
        code, d = s.step(n)
        assert code == 0 and d["dmax"] < p["resid"]


def test_context_that_does_not_fit_fails_cleanly(tc):
    with pytest.raises(tc.TCError, match="hipMalloc"):
        tc.Solver(_params(tc, *HUGE))
    p = _params(tc, 32, 24, 17)
    s = tc.Solver(p)
    _steps_ok(tc, s, p)
    s.close()


def test_batch_instance_that_does_not_fit_is_refused(tc):
    p = _params(tc, 32, 24, 17)
    with tc.Batch() as b:
        i = b.add(p)
        with pytest.raises(tc.TCError, match="hipMalloc"):
            b.add(_params(tc, *HUGE))
        qt, qr, qz = tc.init_perturbed_couette(p, 3, 1e-2)
        s = b.solver(i)
        s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
        s.project()
        assert b.advance(3, graph=True, sync_every=3) == 1
        steps, code, _ = b.status(i)
        assert (steps, code) == (3, 0)
        assert np.isfinite(s.get("qt")).all()
