"""The fused RK3 step applies the pressure correction of substeps 1 and 2
inside the next substep's Helmholtz kernel (TC_DEFER_CORRECTION, on by
default).  Against the per-substep schedule (one correction pass per
substep) the fields agree to rounding, with the tolerance below; where the
deferral does not apply (odd nthm) the two settings are the same code and
agree bit for bit.  Both schedules are checked against the oracle by the
step tests."""
import numpy as np
import pytest

from helpers import small_params

pytestmark = pytest.mark.gpu

FIELDS = ("qt", "qr", "qz", "pr")


def _run(tc, monkeypatch, p, defer, nsteps):
    monkeypatch.setenv("TC_DEFER_CORRECTION", "1" if defer else "0")
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
    ca, da, a = _run(tc, monkeypatch, p, True, 6)
    cb, db, b = _run(tc, monkeypatch, p, False, 6)
    assert ca == cb == 0
    for f in FIELDS:
        err = np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f])
        assert err < 1e-13, (f, err)
    for k in ("dt", "torque_in", "torque_out", "vmax1", "vmax2", "vmax3"):
        assert abs(da[k] - db[k]) <= 1e-12 * abs(db[k]), (k, da[k], db[k])


def test_odd_theta_count_keeps_the_per_substep_schedule(tc, monkeypatch):
    p = small_params(tc, 27, 24, 17, ntst=1000, tout=1e-9)
    _, _, a = _run(tc, monkeypatch, p, True, 3)
    _, _, b = _run(tc, monkeypatch, p, False, 3)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
