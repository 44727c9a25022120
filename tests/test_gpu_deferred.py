"""The fused RK3 step applies the pressure correction of substeps 1 and 2
inside the next substep's Helmholtz kernel (tuning ``defer`` 1 or
TC_DEFER_CORRECTION 1: from 8 Mi grid points up, the default; 2 at any size;
-1, or TC_DEFER_CORRECTION 0: never).  Forced on these small
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


def _run_tuned(tc, p, tuning, nsteps, switch=None):
    s = tc.Solver(p, tuning=tuning)
    qt, qr, qz = tc.init_perturbed_couette(p, 33, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    s.project()
    grew = 0
    for n in range(nsteps):
        if switch and n == switch[0]:
            before = s.device_bytes()
            s.set_tuning(switch[1])
            grew = s.device_bytes() - before
        s.step(n)
    out = {f: s.get(f) for f in FIELDS}
    t = s.tuning()
    s.close()
    return out, dict(t, grew=grew)


def test_tuning_knob_is_the_environment_setting_per_instance(tc, monkeypatch):
    p = small_params(tc, 32, 24, 33, ntst=1000, tout=1e-9)
    _, _, env2 = _run(tc, monkeypatch, p, 2, 3)
    _, _, env0 = _run(tc, monkeypatch, p, 0, 3)
    monkeypatch.delenv("TC_DEFER_CORRECTION")
    always, t2 = _run_tuned(tc, p, {"defer": 2}, 3)
    never, tn = _run_tuned(tc, p, {"defer": -1}, 3)
    default, t1 = _run_tuned(tc, p, None, 3)
    assert (t2["defer"], tn["defer"], t1["defer"]) == (2, -1, 1)
    for f in FIELDS:
        assert np.array_equal(always[f], env2[f]), f
        assert np.array_equal(never[f], env0[f]), f
        assert np.array_equal(default[f], env0[f]), f  # off below 8 Mi points


def test_deferral_switched_on_between_steps(tc, monkeypatch):
    """tc_set_tuning turns the deferral on for a running context (allocating
    the spare pressure array); every step ends with its fields in place, so
    the switch needs no other change of state."""
    monkeypatch.delenv("TC_DEFER_CORRECTION", raising=False)
    p = small_params(tc, 32, 24, 33, ntst=1000, tout=1e-9)
    switched, t = _run_tuned(tc, p, {"defer": -1}, 6, switch=(3, {"defer": 2}))
    plain, _ = _run_tuned(tc, p, {"defer": -1}, 6)
    # the spare pressure array (nr planes) appeared with the switch
    assert t["defer"] == 2 and t["grew"] == 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"], t
    for f in FIELDS:
        err = np.linalg.norm(switched[f] - plain[f]) / np.linalg.norm(plain[f])
        assert err < 1e-13, (f, err)
