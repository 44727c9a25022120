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
