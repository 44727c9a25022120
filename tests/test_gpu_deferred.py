"""The fused RK3 step applies the pressure correction of substeps 1 and 2
inside the next substep's Helmholtz kernel (TC_DEFER_CORRECTION: by default
from 8 Mi grid points up, 2 at any size, 0 never).  Forced on these small
grids, it agrees with the per-substep schedule (one correction pass per
substep) to rounding, with the tolerance below; where the deferral does not
apply (odd nthm, or the default below the size threshold) the settings run
the same code and agree bit for bit.  The step tests check the default
schedule against the oracle at every size they run; the full-size tests
check the deferred one."""
import numpy as np
import pytest

from helpers import small_params

pytestmark = pytest.mark.gpu

FIELDS = ("qt", "qr", "qz", "pr")


def _run(tc, monkeypatch, p, defer, nsteps):
    monkeypatch.setenv("TC_DEFER_CORRECTION", str(defer))
    s = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 33, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    s.project()
    code, last, d = s.run(0, nsteps)
    out = {f: s.get(f) for f in FIELDS}
    s.close()
    return code, d, out


@pytest.mark.parametrize("shape", [(32, 24, 33), (64, 48, 41)])
def test_deferred_correction_matches_per_substep(tc, monkeypatch, shape):
    p = small_params(tc, *shape, ntst=1000, idtv=1, dtmin=1e-3, dtmax=5e-3, limitcfl=0.5,
                     tout=1e-9)
    ca, da, a = _run(tc, monkeypatch, p, 2, 6)
    cb, db, b = _run(tc, monkeypatch, p, 0, 6)
    assert ca == cb == 0
    for f in FIELDS:
        err = np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f])
        assert err < 1e-13, (f, err)
    for k in ("dt", "torque_in", "torque_out", "vmax1", "vmax2", "vmax3"):
        assert abs(da[k] - db[k]) <= 1e-12 * abs(db[k]), (k, da[k], db[k])


def test_odd_theta_count_keeps_the_per_substep_schedule(tc, monkeypatch):
    p = small_params(tc, 27, 24, 17, ntst=1000, tout=1e-9)
    _, _, a = _run(tc, monkeypatch, p, 2, 3)
    _, _, b = _run(tc, monkeypatch, p, 0, 3)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


def test_default_keeps_the_per_substep_schedule_on_small_grids(tc, monkeypatch):
    p = small_params(tc, 32, 24, 33, ntst=1000, tout=1e-9)
    _, _, a = _run(tc, monkeypatch, p, 1, 3)
    _, _, b = _run(tc, monkeypatch, p, 0, 3)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
