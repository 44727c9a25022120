"""tc_run_steps (the main loop inside the library) against tc_step called
once per iteration: identical fields and diagnostics, and the loop ends at
the first error code (555 at NTST) with the last iteration reported."""
import numpy as np
import pytest

from helpers import small_params

pytestmark = pytest.mark.gpu


def _start(tc, p):
    s = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 21, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    s.project()
    return s


def test_run_equals_single_steps(tc):
    p = small_params(tc, 32, 24, 33, ntst=1000, idtv=1, dtmin=1e-3, dtmax=5e-3, limitcfl=0.5,
                     tout=1e-9, sizepert=0.1, omegapert=0.5)
    a, b = _start(tc, p), _start(tc, p)
    for n in range(6):
        code, da = a.step(n)
        assert code == 0
    code, last, db = b.run(0, 6)
    assert code == 0 and last == 5
    assert da == db
    for f in ("qt", "qr", "qz", "pr"):
        assert np.array_equal(a.get(f), b.get(f)), f


def test_run_stops_at_the_first_error_code(tc):
    p = small_params(tc, 32, 24, 33, ntst=4, tout=1e-9)
    s = _start(tc, p)
    code, last, _ = s.run(0, 50)
    assert code == 555 and last == 4
    with pytest.raises(tc.TCError):
        s.run(5, 0)


def test_run_logged_keeps_every_logging_iteration(tc):
    """tc_run_steps_logged returns one row per logging iteration (ntime = 1
    and every TOUT, main.F90:180), each equal to tc_step's diagnostics of
    that iteration."""
    p = small_params(tc, 32, 24, 33, ntst=1000, idtv=0, dtmax=2e-3, dtmin=2e-3, tout=5e-3,
                     sizepert=0.1, omegapert=0.5)
    a, b = _start(tc, p), _start(tc, p)
    want = []
    for n in range(9):
        code, d = a.step(n)
        assert code == 0
        if n == 1 or np.fmod(d["time"], p["tout"]) < d["dt"]:
            want.append((n, d))
    code, last, _, log = b.run_logged(0, 9)
    assert code == 0 and last == 8
    assert [n for n, _ in log] == [n for n, _ in want] and len(want) >= 3
    for (n1, d1), (n2, d2) in zip(log, want):
        assert d1 == d2, (n1, d1, d2)
    code, last, _, log = b.run_logged(9, 5, log_cap=1)
    assert len(log) <= 1


def test_walltime_limit_is_checked_after_the_step(tc):
    """main.F90:208,220: the wall clock is read after the iteration's work,
    so a run started past WALLTIMEMAX still completes its first iteration
    and then reports 334 (and 555 still wins at NTST)."""
    p = small_params(tc, 32, 24, 33, ntst=1000, walltimemax=1.0)
    a, b = _start(tc, p), _start(tc, p)
    code, last, _ = b.run(0, 5, walltime=2.0)
    assert (code, last) == (334, 0)
    a.step(0)
    for f in ("qt", "qr", "qz", "pr"):
        assert np.array_equal(a.get(f), b.get(f)), f
    p2 = dict(p, ntst=0)
    c = _start(tc, p2)
    code, last, _ = c.run(0, 5, walltime=2.0)
    assert (code, last) == (555, 0)
