"""Parity at the production shape, 512 x 512 x 385 (theta x z x r):

- three steps of the production workload itself (seed 2002, variable dt at
  CFL 1.2, oscillating inner cylinder, torque and divergence every step) on
  the device and in the oracle (a few seconds per step on the host cores):
  relative L2 <= 1e-10 on qt, qr, qz and the whole pressure, torques within
  1e-9, divergence below RESID, the same dt sequence;

and through properties that do not need the oracle:

- the fused kernels and the routine-by-routine kernels (two independent
  implementations of the same step) agree to 1e-10 relative L2 after two
  RK3 steps, and each step's maximum divergence stays below RESID;
- an axisymmetric, z-independent Couette start stays axisymmetric and
  z-independent (every r-plane of qt uniform) with q_r and q_z at roundoff;
- the projection is idempotent: projecting a projected field changes it by
  roundoff only, and leaves the divergence at roundoff.
"""
import numpy as np
import pytest

from helpers import compare_fields, rel_l2, start_pair

pytestmark = pytest.mark.gpu


def _production(tc):
    return tc.configs.production_512()


def test_production_512_matches_oracle(tc, oracle):
    p = _production(tc)
    print("full size: production_512, 3 steps, device vs oracle", flush=True)
    dev, orc = start_pair(tc, oracle, p, tc.configs.SEEDS["production_512"],
                          tc.configs.AMPLITUDE["production_512"])
    for n in range(3):
        c1, d1 = dev.step(n)
        c2, d2 = orc.step(n)
        print(f"  step {n}: dt {d1['dt']:.6e} torque {d1['torque_in']:.12f} "
              f"{d1['torque_out']:.12f} dmax {d1['dmax']:.2e}", flush=True)
        assert c1 == c2 == 0, (n, c1, c2)
        assert abs(d1["dt"] / d2["dt"] - 1) < 1e-12, (n, d1["dt"], d2["dt"])
        assert d1["dmax"] < p["resid"]
        assert abs(d1["torque_in"] / d2["torque_in"] - 1) < 1e-9
        assert abs(d1["torque_out"] / d2["torque_out"] - 1) < 1e-9
    e = compare_fields(dev, orc, p["nrm"])
    print(f"  rel L2 {e}", flush=True)
    dev.close()
    orc.close()


def test_fused_and_routine_paths_agree_full_size(tc):
    p = _production(tc)
    print("full size: fused vs routine-by-routine, 2 steps", flush=True)
    qt, qr, qz = tc.init_perturbed_couette(p, tc.configs.SEEDS["production_512"], 1e-2)
    runs = []
    for fused in (True, False):
        dev = tc.Solver(p)
        dev.set_fused(fused)
        dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
        dev.project()
        for n in range(2):
            code, d = dev.step(n)
            assert code == 0, (fused, n, code)
            assert d["dmax"] < p["resid"], (fused, n, d)
        runs.append({f: dev.get(f) for f in ("qt", "qr", "qz", "pr")})
        dev.close()
    del qt, qr, qz
    nrm = p["nrm"]
    for f in ("qt", "qr", "qz", "pr"):
        sl = slice(0, nrm + 1) if f == "qr" else slice(0, nrm)
        e = rel_l2(runs[0][f][sl], runs[1][f][sl])
        print(f"  {f}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-10, (f, e)


def test_axisymmetric_couette_stays_axisymmetric_full_size(tc):
    p = _production(tc)
    print("full size: axisymmetric Couette start, 2 steps", flush=True)
    dev = tc.Solver(p)
    nrm = p["nrm"]
    rm = tc.grid_vector(p, "rm")[:nrm]
    eta = p["rint"] / p["rext"]
    A = (p["virotext"] - eta * eta * p["virotint"]) / (1 - eta * eta)
    B = (p["virotint"] - p["virotext"]) * p["rint"] ** 2 / (1 - eta * eta)
    shape = (nrm + 1, p["nzm"], p["nthm"])
    qt = np.zeros(shape, order="F")
    qt[:nrm] = (A * rm + B / rm)[:, None, None]
    dev.set("qt", qt)
    del qt
    assert dev.CheckDivergence() == 0.0
    for n in range(2):
        code, d = dev.step(n)
        assert code == 0
    qt, qr, qz = dev.get("qt")[:nrm], dev.get("qr"), dev.get("qz")[:nrm]
    scale = np.abs(qt).max()
    spread = (qt.max(axis=(1, 2)) - qt.min(axis=(1, 2))).max()
    print(f"  qt plane spread {spread:.2e}, max|qr| {np.abs(qr).max():.2e}, "
          f"max|qz| {np.abs(qz).max():.2e}, scale {scale:.3f}", flush=True)
    assert spread <= 1e-13 * scale
    assert np.abs(qr).max() <= 1e-12 * scale
    assert np.abs(qz).max() <= 1e-12 * scale
    dev.close()


def test_projection_idempotent_full_size(tc):
    p = _production(tc)
    print("full size: projection idempotence", flush=True)
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 99, 1e-2)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    del qt, qr, qz
    assert dev.CheckDivergence() > 1e-3
    dev.project()
    d1 = dev.CheckDivergence()
    a = {f: dev.get(f) for f in ("qt", "qr", "qz")}
    dev.project()
    d2 = dev.CheckDivergence()
    print(f"  max div after projections: {d1:.2e}, {d2:.2e}", flush=True)
    assert d1 < 1e-11 and d2 < 1e-11
    for f in ("qt", "qr", "qz"):
        e = np.abs(dev.get(f) - a[f]).max() / np.abs(a["qt"]).max()
        assert e <= 1e-13, (f, e)
    dev.close()


def _stats(dev):
    vmax, rms = dev.CalcVelocityStats()
    t_in, t_out = dev.CalcTorque()
    return np.concatenate([vmax, rms, [t_in, t_out, dev.CalcMaxCFL()]])


def test_large_1024_fused_and_routine_paths_agree(tc):
    """1024 x 1024 x 769 (eta 0.909, counter-rotating outer cylinder), the
    largest workload: projection to roundoff, two RK3 steps below RESID on
    each kernel set, and the two kernel sets agree on the velocity maxima,
    mean squares, wall torques and CFL (device reductions, so the 6.5 GB
    fields never leave the device)."""
    p = tc.configs.large_1024()
    print("1024x1024x769: fused vs routine-by-routine, 2 steps", flush=True)
    qt, qr, qz = tc.init_perturbed_couette(p, tc.configs.SEEDS["large_1024"],
                                           tc.configs.AMPLITUDE["large_1024"])
    out = []
    for fused in (True, False):
        dev = tc.Solver(p)
        dev.set_fused(fused)
        dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
        dev.project()
        assert dev.CheckDivergence() < 1e-10
        for n in range(2):
            code, d = dev.step(n)
            assert code == 0 and d["dmax"] < p["resid"], (fused, n, code, d)
        out.append(_stats(dev))
        print(f"  fused={fused}: {out[-1]}", flush=True)
        dev.close()
    assert np.allclose(out[0], out[1], rtol=1e-11, atol=0), (out[0], out[1])
