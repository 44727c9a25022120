"""Whole-step parity: the main.F90 loop on the device against the oracle.

Bar (the workload contract): relative L2 <= 1e-10 on qt, qr, qz and on the
pressure (modulo the theta-z mean profile that the singular (0,0) Poisson
mode leaves free in the reference), max divergence below RESID, wall torques
within 1e-9 relative.
"""
import numpy as np
import pytest

from helpers import pressure_modulo_gauge, rel_l2

pytestmark = pytest.mark.gpu


def _start(tc, oracle, p, seed, amp):
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    dev.project()
    orc = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        orc.set(n, dev.get(n))
    return dev, orc


def _compare(dev, orc, nrm, tol=1e-10):
    out = {}
    for n, sl in (("qt", slice(0, nrm)), ("qr", slice(0, nrm + 1)), ("qz", slice(0, nrm))):
        out[n] = rel_l2(dev.get(n)[sl], orc.get(n)[sl])
    a, b = dev.get("pr")[:nrm], orc.get("pr")[:nrm]
    out["pr"] = np.linalg.norm(pressure_modulo_gauge(a, b)) / np.linalg.norm(b - b.mean(axis=(1, 2), keepdims=True))
    for n, e in out.items():
        assert e <= tol, f"{n}: rel L2 {e:.3e} ({out})"
    return out


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


def test_small_64_twenty_steps(tc, oracle):
    p = tc.configs.small_64()
    dev, orc = _start(tc, oracle, p, tc.configs.SEEDS["small_64"], 1e-2)
    for n in range(20):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        assert c1 == c2 == 0
    e = _compare(dev, orc, p["nrm"])
    ti, to = dev.CalcTorque()
    ri, ro = orc.torque()
    assert abs(ti / ri - 1) < 1e-9 and abs(to / ro - 1) < 1e-9
    print("small_64 20 steps", e)
