"""Checkpoint / restart: a run interrupted by tc_write_restart and continued
from tc_read_restart in a fresh context is bit-identical to the
uninterrupted run (fields, pressure and the diagnostics of every step), with
variable dt and an oscillating inner cylinder; a file of another grid is
rejected."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(tc):
    p = tc.configs.scaled(tc.configs.production_512(), 32, 24, 33)
    p.update(sizepert=0.1, omegapert=0.5, ntst=1000)
    return p


def test_restart_continues_bit_for_bit(tc, tmp_path):
    p = _params(tc)
    qt, qr, qz = tc.init_perturbed_couette(p, 41, 1e-2)
    a = tc.Solver(p)
    a.set("qt", qt); a.set("qr", qr); a.set("qz", qz)
    a.project()
    diags = []
    path = os.path.join(tmp_path, "ckpt.bin")
    for n in range(7):
        if n == 4:
            a.write_restart(path, n)
        diags.append(a.step(n))
    b = tc.Solver(p)
    n0 = b.read_restart(path)
    assert n0 == 4
    for n in range(n0, 7):
        code, d = b.step(n)
        assert code == diags[n][0]
        assert d == diags[n][1], (n, d, diags[n][1])
    for f in ("qt", "qr", "qz", "pr"):
        assert np.array_equal(a.get(f), b.get(f)), f
    other = tc.Solver(tc.configs.scaled(p, 32, 24, 25))
    with pytest.raises(tc.TCError):
        other.read_restart(path)
    for s in (a, b, other):
        s.close()
