"""Statistics on the device against the oracle: CalcStats (with the vorticity
of CalcVorticity) and CalcDissipation, routine by routine and inside the
main-loop step once time > TSTA, plus a larger shape checked against numpy
theta-means of the downloaded fields.

Bar: each accumulator within 1e-11 of the oracle's relative to its largest
magnitude (the device sums each theta line with a fixed butterfly, the
oracle point by point, so the roundings differ), nute within 1e-11
relative; in the step loop, where the fields themselves differ at the
1e-13 level, 1e-9.
"""
import numpy as np
import pytest

from helpers import load_both, random_state, small_params

pytestmark = pytest.mark.gpu


def _cmp_stats(dev, orc, tol):
    a, na = dev.stats()
    b, nb = orc.stats()
    assert na == nb
    for name in b:
        scale = np.abs(b[name]).max()
        err = np.abs(a[name] - b[name]).max() / (scale if scale > 0 else 1.0)
        assert err <= tol, f"{name}: {err:.3e}"


@pytest.mark.parametrize("shape", [(32, 24, 33), (48, 40, 17), (100, 18, 12)])
def test_stats_and_dissipation_routines(tc, oracle, shape):
    nthm, nzm, nrm = shape
    p = small_params(tc, nthm, nzm, nrm, virotext=-0.2)
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    load_both(dev, orc, random_state(p, 31))
    for _ in range(2):
        dev.CalcStats()
        orc.calc_stats()
    nd = dev.CalcDissipation()
    no = orc.calc_dissipation()
    assert abs(nd - no) <= 1e-11 * abs(no), (nd, no)
    assert dev.last_dissipation() == nd
    _cmp_stats(dev, orc, 1e-11)
    # a dissipation-only call adds to disste and leaves the rest alone
    before, n0 = dev.stats()
    dev.CalcDissipation()
    after, n1 = dev.stats()
    assert n1 == n0 == 2
    for name in before:
        if name == "disste":
            np.testing.assert_allclose(after[name], 2 * before[name], rtol=1e-13)
        else:
            assert np.array_equal(after[name], before[name]), name
    # set_stats round trip
    dev.set_stats(before, 7)
    again, n2 = dev.stats()
    assert n2 == 7 and all(np.array_equal(again[k], before[k]) for k in before)


def test_stats_in_the_step_loop(tc, oracle):
    p = small_params(tc, 32, 24, 33, tsta=0.004, tout=1e-9, ntst=6, virotext=-0.2,
                     idtv=1, dtmin=2e-3, dtmax=4e-3, limitcfl=0.5)
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    load_both(dev, orc, random_state(p, 5, pressure=False, history=False))
    dev.project()
    for n in ("qt", "qr", "qz"):
        orc.set(n, dev.get(n))
    counts = []
    for n in range(p["ntst"] + 1):
        cd, _ = dev.step(n)
        co, _ = orc.step(n)
        assert cd == co
        counts.append(dev.stats()[1])
    assert counts[0] == 0 and counts[-1] >= 3, counts
    _cmp_stats(dev, orc, 1e-9)
    assert abs(dev.last_dissipation() - orc.last_nute()) <= 1e-9 * abs(orc.last_nute())


def test_stats_match_numpy_theta_means(tc):
    """Size-independent facts at a larger, ragged shape: the means of qt, qz
    and pr and the rms of qt are theta-means of the downloaded fields."""
    nthm, nzm, nrm = 200, 54, 49
    p = small_params(tc, nthm, nzm, nrm)
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 8, 0.05)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    dev.project()
    rng = np.random.default_rng(3)
    pr = np.zeros(dev.shape, order="F")
    pr[:nrm] = rng.uniform(-1, 1, (nrm, nzm, nthm))
    dev.set("pr", pr)
    dev.CalcStats()
    st, n = dev.stats()
    assert n == 1
    qt, qz = dev.get("qt")[:nrm], dev.get("qz")[:nrm]
    for name, want in (("vt_mean", qt.mean(axis=2)), ("vz_mean", qz.mean(axis=2)),
                       ("pr_mean", pr[:nrm].mean(axis=2)), ("vt_rms", (qt * qt).mean(axis=2))):
        err = np.abs(st[name] - want).max() / np.abs(want).max()
        assert err < 1e-13, (name, err)


@pytest.mark.parametrize("shape", [(32, 24, 33), (100, 18, 12)])
def test_slab_dumps(tc, oracle, shape):
    nthm, nzm, nrm = shape
    p = small_params(tc, nthm, nzm, nrm)
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    load_both(dev, orc, random_state(p, 17))
    for k in (1, nrm // 2, nrm):
        for a, b in zip(dev.DumpRCut(k), orc.dump_rcut(k)):
            np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-15)
    for a, b in zip(dev.DumpThCut(), orc.dump_thcut()):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-15)
    for a, b in zip(dev.DumpStreamwiseAvg(), orc.dump_streamwise_avg()):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)
    with pytest.raises(tc.TCError):
        dev.DumpRCut(nrm + 1)
