"""Whole-step parity: the main.F90 loop on the device against the oracle.

Bar (the workload contract): relative L2 <= 1e-10 on qt, qr, qz and on the
whole pressure (device and oracle fix the gauge of the singular (0,0)
Poisson mode alike), max divergence below RESID, wall torques within 1e-9
relative.
"""
import numpy as np
import pytest

from helpers import compare_fields, start_pair

pytestmark = pytest.mark.gpu


def _start(tc, oracle, p, seed, amp):
    return start_pair(tc, oracle, p, seed, amp)


def _compare(dev, orc, nrm, tol=1e-10):
    return compare_fields(dev, orc, nrm, tol)


@pytest.mark.parametrize("fused", [True, False])
def test_projection_is_divergence_free(tc, oracle, fused):
    p = tc.configs.scaled(tc.configs.small_64(), 48, 32, 33)
    dev = tc.Solver(p)
    dev.set_fused(fused)
    qt, qr, qz = tc.init_perturbed_couette(p, 5, 1e-2)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    assert dev.CheckDivergence() > 1e-3
    dev.project()
    assert dev.CheckDivergence() < 1e-12


@pytest.mark.parametrize("fused", [True, False])
def test_steps_match_oracle_small(tc, oracle, fused):
    p = tc.configs.scaled(tc.configs.small_64(), 32, 24, 33)
    p.update(sizepert=0.1, omegapert=0.5, virotext=-0.3)
    dev, orc = _start(tc, oracle, p, 7, 1e-2)
    dev.set_fused(fused)
    for n in range(8):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2
        assert d1["dt"] == d2["dt"]
        if d2["dmax"] > 0:
            assert d1["dmax"] < p["resid"]
            assert abs(d1["torque_in"] / d2["torque_in"] - 1) < 1e-9
            assert abs(d1["torque_out"] / d2["torque_out"] - 1) < 1e-9
    _compare(dev, orc, p["nrm"])


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("shape", [(40, 45, 17), (27, 20, 13), (45, 27, 9), (16, 12, 3), (8, 4, 5)])
def test_odd_shapes_match_oracle(tc, oracle, shape, fused):
    """Odd theta or z sizes (3- and 5-smooth): theta-FFT row pairs that
    straddle two r-planes and a last unpaired row (odd nzm*nrm), and, for odd
    nthm, the paths without theta-paired corrections or the fused
    end-of-step pass; the smallest radial size (nrm = 3: one Poisson
    partition per line) and a theta size below one tile width."""
    p = tc.configs.scaled(tc.configs.small_64(), *shape)
    p.update(sizepert=0.1, omegapert=0.5, virotext=-0.3)
    dev, orc = _start(tc, oracle, p, 23, 1e-2)
    dev.set_fused(fused)
    for n in range(5):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2
        assert d1["dt"] == d2["dt"]
        if d2["dmax"] > 0:
            assert d1["dmax"] < p["resid"]
            assert abs(d1["torque_in"] / d2["torque_in"] - 1) < 1e-9
            assert abs(d1["torque_out"] / d2["torque_out"] - 1) < 1e-9
    _compare(dev, orc, p["nrm"])


@pytest.mark.parametrize("defer", ["0", "2"])
@pytest.mark.parametrize("lpt", ["1", "2"])
@pytest.mark.parametrize("masks", [False, True])
def test_fused_variants_match_oracle(tc, oracle, monkeypatch, lpt, masks, defer):
    """Both tilings of the fused Helmholtz kernel (one or two radial lines per
    thread), with uniform and with mixed slip/no-slip wall masks, with the
    pressure correction after every substep and deferred into the next
    substep's Helmholtz kernel."""
    monkeypatch.setenv("TC_HELM_LPT", lpt)
    monkeypatch.setenv("TC_DEFER_CORRECTION", defer)
    p = tc.configs.scaled(tc.configs.small_64(), 96, 40, 29)
    p.update(sizepert=0.05, omegapert=0.3, virotext=0.2, romeg=0.1)
    if masks:
        # free-slip strips on the inner cylinder from SlipBCs
        p.update(minstrt=2, minstrz=3, multmin=1)
    dev, orc = _start(tc, oracle, p, 17, 1e-2)
    for n in range(4):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2
    if masks:
        assert dev.masks()[0].mean() < 1.0
    _compare(dev, orc, p["nrm"])


def test_variable_dt_steps_match_oracle(tc, oracle):
    p = tc.configs.scaled(tc.configs.production_512(), 48, 40, 33)
    p.update(ntst=5)
    dev, orc = _start(tc, oracle, p, 9, 1e-2)
    codes = []
    for n in range(6):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        codes.append(c1)
        assert c1 == c2
        assert abs(d1["dt"] / d2["dt"] - 1) < 1e-12
        assert d1["dmax"] < p["resid"]
        assert abs(d1["torque_in"] / d2["torque_in"] - 1) < 1e-9
        assert abs(d1["torque_out"] / d2["torque_out"] - 1) < 1e-9
    assert codes[-1] == 555
    _compare(dev, orc, p["nrm"])


@pytest.mark.parametrize("tout", [1e30, 0.004])
def test_variable_dt_without_logging_matches_oracle(tc, oracle, tout):
    """Variable dt where most steps do not log (TOUT = 1e30: only ntime 1;
    0.004: every few steps): the device takes each non-logging step's CFL
    from the previous step's fused last correction pass instead of a CFL
    pass of its own (CalcMaxCFL of main.F90:156 on the same fields); the dt
    sequence and the fields match the oracle, which evaluates CalcMaxCFL at
    every step start."""
    p = tc.configs.scaled(tc.configs.production_512(), 48, 40, 33)
    p.update(ntst=1000, tout=tout)
    dev, orc = _start(tc, oracle, p, 17, 1e-2)
    logged = 0
    for n in range(12):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2 == 0, (n, c1, c2)
        assert abs(d1["dt"] / d2["dt"] - 1) < 1e-12, (n, d1["dt"], d2["dt"])
        logged += d1["vmax1"] != 0.0
    assert logged < 12
    assert dev.CheckDivergence() < p["resid"]
    _compare(dev, orc, p["nrm"])


def test_small_64_full_run(tc, oracle):
    """The 64x64x65 workload as specified: RK3 at fixed dt 1e-3, NTST = 199,
    i.e. 200 steps ending with code 555 (main.F90:152,223), through the
    library's own main loop on the device and step by step in the oracle."""
    p = tc.configs.small_64()
    dev, orc = _start(tc, oracle, p, tc.configs.SEEDS["small_64"], 1e-2)
    print("small_64: 200 RK3 steps, device main loop vs oracle", flush=True)
    code, last, d1 = dev.run(0, 1000)
    assert (code, last) == (555, 199), (code, last)
    for n in range(200):
        c2, d2 = orc.step(n)
        assert c2 == (555 if n == 199 else 0), (n, c2)
    assert d1["time"] == d2["time"] and d1["dt"] == d2["dt"] == 1e-3
    e = _compare(dev, orc, p["nrm"])
    ti, to = dev.CalcTorque()
    ri, ro = orc.torque()
    assert abs(ti / ri - 1) < 1e-9 and abs(to / ro - 1) < 1e-9
    assert dev.CheckDivergence() < p["resid"]
    print("small_64 200 steps", e, flush=True)


@pytest.mark.parametrize("fused", [True, False])
def test_ab2_steps_match_oracle(tc, oracle, fused):
    """NSST = 1: one Adams-Bashforth 2 substep per step (gamma 3/2, rho -1/2,
    InitTimeMarchScheme.F90:29-36), through both kernel sets, with the
    inner cylinder oscillating and variable dt."""
    p = tc.configs.scaled(tc.configs.production_512(), 40, 32, 29)
    p.update(nsst=1, ntst=1000, sizepert=0.1, omegapert=0.4, virotext=-0.2, dtmax=2e-3)
    dev, orc = _start(tc, oracle, p, 23, 1e-2)
    dev.set_fused(fused)
    for n in range(8):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2 == 0
        assert abs(d1["dt"] / d2["dt"] - 1) < 1e-12
        assert d1["dmax"] < p["resid"]
        assert abs(d1["torque_in"] / d2["torque_in"] - 1) < 1e-9
        assert abs(d1["torque_out"] / d2["torque_out"] - 1) < 1e-9
    _compare(dev, orc, p["nrm"])
