"""Mathematical pins of the oracle (CPU only).

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, an index off by one, a transposed coefficient) fails it:
  - the explicit terms converge at second order to the continuous
    cylindrical Navier-Stokes terms they discretise (analytic fields);
  - the radial operators of the Helmholtz solves converge to theirs;
  - a pure radial-diffusion Bessel mode decays at the analytic rate through
    full RK3 / Crank-Nicolson steps (Helmholtz assembly, wall rows, LU
    solve, RK coefficients, projection);
  - divergence of the pressure-correction gradient inverts the Poisson solve;
  - closed forms: solid-body rotation, CorrectPressure on a constant, the
    Couette torque, the untouched wall history of ImplicitAndUpdateQZ;
  - the main-loop error codes forced from bou.in values.
"""
import numpy as np
import pytest
from scipy.optimize import brentq
from scipy.special import jv, yv

import analytic as A


def params(tc, nthm, nzm, nrm, **kw):
    p = tc.configs.small_64()
    p.update(nthm=nthm, nzm=nzm, nrm=nrm, ren=10.0, romeg=0.7, rint=1.0, rext=1.5,
             lamb=4.0, alx3=1.0, istr=4, str=1.5)
    p.update(kw)
    return p


def explicit_errors(tc, oracle, n):
    p = params(tc, n, n, n)
    o = oracle.Oracle(p)
    F = A.Field(p)
    qt, qr, qz = A.sample_state(o, F)
    o.set("qt", qt); o.set("qr", qr); o.set("qz", qz)
    o.call("explicit_terms_qt"); o.call("explicit_terms_qr"); o.call("explicit_terms_qz")
    rc, rm, tc_, tm, zc, zm = A.staggered(o)
    ks = slice(2, n - 2)
    errs = {}
    R, Z, T = np.meshgrid(rm[ks], zm, tc_, indexing="ij")
    ref = A.cont_qt(F, p["ren"], p["romeg"], R, T, Z)
    errs["dq"] = np.abs(o.get("dq")[ks] - ref).max() / np.abs(ref).max()
    R, Z, T = np.meshgrid(rc[ks], zm, tm, indexing="ij")
    ref = A.cont_qr(F, p["ren"], p["romeg"], R, T, Z)
    errs["dph"] = np.abs(o.get("dph")[ks] - ref).max() / np.abs(ref).max()
    R, Z, T = np.meshgrid(rm[ks], zc, tm, indexing="ij")
    ref = A.cont_qz(F, p["ren"], R, T, Z)
    errs["qcap"] = np.abs(o.get("qcap")[ks] - ref).max() / np.abs(ref).max()
    return errs


def test_explicit_terms_converge_second_order(tc, oracle):
    e1 = explicit_errors(tc, oracle, 24)
    e2 = explicit_errors(tc, oracle, 48)
    for k in e1:
        ratio = e2[k] / e1[k]
        assert e1[k] < 0.2 and ratio < 0.35, (k, e1[k], e2[k], ratio)


def test_explicit_terms_solid_body(tc, oracle):
    """qt = Omega rm, qr = qz = 0, no rotation frame: dq = qcap = 0 and dph is
    the centrifugal term ((qt(k) + qt(k-1))/2)^2 (ExplicitTermsQR.F90:58-59)."""
    p = params(tc, 16, 12, 17, romeg=0.0)
    o = oracle.Oracle(p)
    rm = o.vec("rm")
    om = 1.7
    qt = np.zeros(o.shape, order="F")
    qt[: o.nrm] = (om * rm[: o.nrm])[:, None, None]
    o.set("qt", qt)
    o.set("qr", np.zeros(o.shape)); o.set("qz", np.zeros(o.shape))
    o.call("explicit_terms_qt"); o.call("explicit_terms_qr"); o.call("explicit_terms_qz")
    assert np.abs(o.get("dq")[: o.nrm]).max() == 0.0
    assert np.abs(o.get("qcap")[: o.nrm]).max() == 0.0
    cent = (0.5 * om * (rm[1:o.nrm] + rm[: o.nrm - 1])) ** 2
    np.testing.assert_allclose(o.get("dph")[1:o.nrm], np.broadcast_to(cent[:, None, None], (o.nrm - 1, o.nzm, o.nthm)), rtol=1e-14)


def radial_operator_errors(tc, oracle, n):
    p = params(tc, 4, 4, n)
    o = oracle.Oracle(p)
    rc, rm = o.vec("rc"), o.vec("rm")[:n]
    ri, ro = p["rint"], p["rext"]
    f = lambda r: np.sin(np.pi * (r - ri) / (ro - ri)) + 0.3 * r * r
    fp = lambda r: A.d1(f, r, A.H1)
    fpp = lambda r: A.d2(f, r, A.H2)
    out = {}
    k = np.arange(1, n - 1)  # rows 2..nrm-1 (0-based 1..n-2)
    for name, lhs in (("A1", fpp(rm) + fp(rm) / rm - f(rm) / rm ** 2), ("A3", fpp(rm) + fp(rm) / rm)):
        ap, ac, am = (o.vec(s + name[1] + "j")[:n] for s in ("ap", "ac", "am"))
        app = ap[k] * f(rm[k + 1]) + ac[k] * f(rm[k]) + am[k] * f(rm[k - 1])
        out[name] = np.abs(app - lhs[k]).max() / np.abs(lhs[k]).max()
    kk = np.arange(1, n - 1)  # rc rows 2..nrm (0-based)
    lhs = fpp(rc) - fp(rc) / rc
    for name, sfx in (("A2", "2j"), ("A2e", "2je")):
        ap, ac, am = (o.vec(s + sfx)[: n + 1] for s in ("ap", "ac", "am"))
        app = ap[kk] * f(rc[kk + 1]) + ac[kk] * f(rc[kk]) + am[kk] * f(rc[kk - 1])
        out[name] = np.abs(app - lhs[kk]).max() / np.abs(lhs[kk]).max()
    return out


def test_radial_operators_converge_second_order(tc, oracle):
    """CreateGrid.F90:194-322 coefficients: A1 = d2/dr2 + (1/r) d/dr - 1/r^2,
    A3 = d2/dr2 + (1/r) d/dr (at rm); A2, A2e = d2/dr2 - (1/r) d/dr on
    q_r = r v_r (at rc)."""
    e1 = radial_operator_errors(tc, oracle, 32)
    e2 = radial_operator_errors(tc, oracle, 64)
    for k in e1:
        assert e2[k] / e1[k] < 0.35 and e1[k] < 0.05, (k, e1[k], e2[k])


def bessel_alpha(order, ri, ro):
    g = lambda a: jv(order, a * ri) * yv(order, a * ro) - jv(order, a * ro) * yv(order, a * ri)
    a = np.linspace(0.5, 40, 4000)
    v = g(a)
    i = np.nonzero(np.sign(v[:-1]) != np.sign(v[1:]))[0][0]
    return brentq(g, a[i], a[i + 1])


def decay_rate(tc, oracle, comp, nrm):
    # Re and dt keep beta * max eigenvalue of the Laplacian in the range where
    # the reference's incremental pressure update (be = beta,
    # CorrectPressure.F90:23) is stable
    ren, ri, ro = 100.0, 1.0, 1.5
    p = params(tc, 4, 4, nrm, ren=ren, romeg=0.0, idtv=0, dtmin=1e-3, dtmax=1e-3, istr=0,
               virotint=0.0, virotext=0.0, limitcfl=1e30, tout=1e30)
    o = oracle.Oracle(p)
    order = 1 if comp == "qt" else 0
    al = bessel_alpha(order, ri, ro)
    rm = o.vec("rm")[:nrm]
    prof = jv(order, al * rm) * yv(order, al * ri) - yv(order, al * rm) * jv(order, al * ri)
    q = np.zeros(o.shape, order="F")
    q[:nrm] = prof[:, None, None]
    o.set(comp, q)
    norms, times = [], []
    for n in range(200):
        o.step(n)
        if n in (99, 199):
            norms.append(np.linalg.norm(o.get(comp)))
            times.append(o.scalars()["time"])
    qr = np.abs(o.get("qr")).max()
    rate = np.log(norms[0] / norms[1]) / (times[1] - times[0])
    return rate, al * al / ren, qr


@pytest.mark.parametrize("comp", ["qz", "qt"])
def test_bessel_mode_decay(tc, oracle, comp):
    """Axisymmetric, z-independent v_z (J0/Y0) or v_theta (J1/Y1) between
    no-slip walls decays as exp(-alpha^2 t / Re) through full time steps;
    the centrifugal forcing of the v_theta mode is removed by the projection."""
    r1, exact, qr1 = decay_rate(tc, oracle, comp, 16)
    r2, _, qr2 = decay_rate(tc, oracle, comp, 32)
    e1, e2 = abs(r1 / exact - 1), abs(r2 / exact - 1)
    assert e1 < 0.05 and e2 / e1 < 0.35, (comp, r1, r2, exact, e1, e2)
    assert max(qr1, qr2) < 1e-10


def test_poisson_inverts_divergence_of_gradient(tc, oracle):
    """SolvePressureCorrection is the inverse of D(G .), with D the divergence
    of CalcLocalDivergence.F90:33-35 and G the gradient of CorrectVelocity.F90
    :35-40, except for the singular (0,0) mode at its last row (gauge)."""
    for shape in ((16, 12, 17), (30, 18, 21), (24, 20, 33)):
        p = params(tc, *shape)
        o = oracle.Oracle(p)
        nrm = o.nrm
        rng = np.random.default_rng(5)
        rhs = np.zeros(o.shape, order="F")
        rhs[:nrm] = rng.normal(size=(nrm, o.nzm, o.nthm))
        o.set("dph", rhs)
        o.call("solve_pressure_correction")
        phi = o.get("dph")
        for f in ("qt", "qr", "qz"):
            o.set(f, np.zeros(o.shape))
        o.set_time(0.0, 1.0)
        o.set_substep(1)
        al = o.scalars()["al"]
        o.set_time(0.0, 1.0 / al)  # al*dt = 1
        o.set("dph", phi)
        o.call("correct_velocity")
        for f in ("qt", "qr", "qz"):  # TimeMarcher.F90:51-53
            o.update_halo(f)
        o.call("calc_local_divergence")
        lphi = -o.get("dph")[:nrm]
        d = rhs[:nrm] - lphi
        fluct = d - d.mean(axis=(1, 2), keepdims=True)
        assert np.abs(fluct).max() < 1e-9 * np.abs(rhs).max(), np.abs(fluct).max()
        mean = d.mean(axis=(1, 2))
        assert np.abs(mean[:-1]).max() < 1e-9 * np.abs(rhs).max()


def test_correct_pressure_constant_correction(tc, oracle):
    """CorrectPressure.F90:40-46 uses apph for both radial neighbours and
    be = beta: a constant phi = c gives pr = c (1 - beta (apph - amph))."""
    p = params(tc, 8, 8, 13)
    o = oracle.Oracle(p)
    c = 0.37
    dph = np.zeros(o.shape, order="F")
    dph[: o.nrm] = c
    o.set("dph", dph)
    o.set("pr", np.zeros(o.shape))
    o.set_time(0.0, 3e-3)
    o.begin_step()
    beta = o.scalars()["beta"]
    assert beta == 3e-3 / p["ren"] * 0.5
    o.call("correct_pressure")
    apph, amph = o.vec("apph")[: o.nrm], o.vec("amph")[: o.nrm]
    expect = c * (1 - beta * (apph - amph))
    np.testing.assert_allclose(o.get("pr")[: o.nrm], np.broadcast_to(expect[:, None, None], (o.nrm, 8, 8)), rtol=1e-13)
    assert abs(expect[0] - c) > 1e-6  # the profile is not flat


def test_qz_wall_history_is_overwritten(tc, oracle):
    """ImplicitAndUpdateQZ.F90:67-113 assembles row nrm twice; the second
    pass sees ru3(nrm) already replaced by qcap(nrm), so the stored history
    of that row never reaches the solve, while every other row's does."""
    p = params(tc, 8, 6, 11)
    base = None
    results = {}
    for case in ("none", "wall", "inner"):
        o = oracle.Oracle(p)
        ru3 = np.zeros(o.shape, order="F")
        if case == "wall":
            ru3[o.nrm - 1] = 1.0
        elif case == "inner":
            ru3[o.nrm - 2] = 1.0
        o.set("ru3", ru3)
        o.set_time(0.0, 1e-3)
        o.begin_step()
        o.set_substep(2)  # ro != 0
        o.call("implicit_and_update_qz")
        results[case] = o.get("qz")
    assert np.array_equal(results["wall"], results["none"])
    assert not np.array_equal(results["inner"], results["none"])


@pytest.mark.parametrize("nrm", [33])
def test_couette_torque_converges(tc, oracle, nrm):
    """CalcTorque.F90:26-55 on the laminar Couette profile v = A r + B / r
    with the wall velocities VIROTINT, VIROTEXT of the boundary conditions
    (ImplicitAndUpdateQT.F90:57,105) tends to 1.  The wall gradient is a
    one-sided difference weighted with rc(2)^3 (rc(nrm)^3), so the
    convergence is first order."""
    errs = []
    for n in (nrm, 2 * nrm - 1):
        p = params(tc, 4, 4, n, virotint=1.0, virotext=-0.3, istr=0)
        o = oracle.Oracle(p)
        ri, ro, vi, vo = p["rint"], p["rext"], p["virotint"], p["virotext"]
        Al = (vo * ro - vi * ri) / (ro * ro - ri * ri)
        Bl = (vi - Al * ri) * ri
        rm = o.vec("rm")[:n]
        qt = np.zeros(o.shape, order="F")
        qt[:n] = (Al * rm + Bl / rm)[:, None, None]
        o.set("qt", qt)
        ti, to = o.torque()
        errs.append(max(abs(ti - 1), abs(to - 1)))
    assert errs[0] < 0.05 and 0.4 < errs[1] / errs[0] < 0.6, errs


def _step_codes(oracle, p, n, walltime=0.0):
    o = oracle.Oracle(p)
    rng = np.random.default_rng(1)
    q = np.zeros(o.shape, order="F")
    q[1 : o.nrm] = 1e-2 * rng.uniform(-1, 1, (o.nrm - 1, o.nzm, o.nthm))
    o.set("qr", q)
    o.call("project")
    codes = []
    for k in range(n):
        c, _ = o.step(k, walltime)
        codes.append(c)
        if c:
            break
    return codes


def test_main_loop_error_codes(tc, oracle):
    """main.F90:158-223 exit codes, each forced from one bou.in value."""
    base = params(tc, 8, 8, 9, ren=100.0, romeg=0.0)
    base.update(tout=1e30, ntst=50, tmax=1e30)
    assert _step_codes(oracle, dict(base, ntst=2), 5) == [0, 0, 555]
    assert _step_codes(oracle, dict(base, resid=1e-300), 5) == [0, 169]
    assert _step_codes(oracle, dict(base, tmax=1e-6), 5) == [333]
    assert _step_codes(oracle, dict(base, limitvel=1e-9), 5) == [0, 266]
    assert _step_codes(oracle, dict(base, idtv=0, dtmax=1e-3, limitcfl=1e-9), 5) == [165]
    assert _step_codes(oracle, dict(base, idtv=1, dtmin=2e-3, dtmax=1e-2, limitcfl=1e-9), 5) == [0, 0, 166]
    assert _step_codes(oracle, dict(base, walltimemax=1.0), 5, walltime=2.0) == [334]


def _reduction_oracle(tc, oracle, nthm=12, nzm=10, nrm=9):
    p = params(tc, nthm, nzm, nrm, alx3=1.7, lamb=3.0)
    return p, oracle.Oracle(p)


def test_max_cfl_closed_form(tc, oracle):
    """CalcMaxCFL.F90:22-39.  Solid body qt = Omega rm gives the same
    azimuthal term Omega dxt at every radius; a uniform qz adds W dxz; the
    zero field returns the 1e-8 floor."""
    p, o = _reduction_oracle(tc, oracle)
    shape = o.shape
    assert o.max_cfl() == 1e-8
    rm = o.vec("rm")
    Om, W = 0.8, -0.3
    qt = np.zeros(shape, order="F")
    qt[: o.nrm] = Om * rm[: o.nrm][:, None, None]
    qz = np.full(shape, W, order="F")
    o.set("qt", qt)
    o.set("qz", qz)
    dxt = p["nthm"] * p["lamb"] / (2 * np.pi)
    dxz = p["nzm"] / p["alx3"]
    assert abs(o.max_cfl() - (Om * dxt + abs(W) * dxz)) <= 1e-13 * (Om * dxt + abs(W) * dxz)
    o.close()


def test_check_divergence_closed_forms(tc, oracle):
    """CheckDivergence.F90:21-40.  qt = f(z), qz = g(theta), qr = 0 is
    discretely solenoidal (exact zero); a one-point spike in qz gives
    max |div| = dxz; a spike in qt at theta index 0 on every (r, z) gives
    dxt / rm(1); a spike qr(2) = 1 gives udx2m(k)/rm(k) at k = 1, 2."""
    p, o = _reduction_oracle(tc, oracle)
    rng = np.random.default_rng(3)
    shape = o.shape
    dxt = p["nthm"] * p["lamb"] / (2 * np.pi)
    dxz = p["nzm"] / p["alx3"]
    qt = np.zeros(shape, order="F")
    qt[: o.nrm] = rng.uniform(-1, 1, o.nzm)[None, :, None]
    qz = np.zeros(shape, order="F")
    qz[: o.nrm] = rng.uniform(-1, 1, o.nthm)[None, None, :]
    o.set("qt", qt)
    o.set("qz", qz)
    o.set("qr", np.zeros(shape))
    assert o.max_div() == 0.0
    o.set("qt", np.zeros(shape))
    qz = np.zeros(shape, order="F")
    qz[3, 4, 5] = 1.0
    o.set("qz", qz)
    assert abs(o.max_div() - dxz) <= 1e-14 * dxz
    o.set("qz", np.zeros(shape))
    qt = np.zeros(shape, order="F")
    qt[: o.nrm, :, 0] = 1.0
    o.set("qt", qt)
    rm = o.vec("rm")
    assert abs(o.max_div() - dxt / rm[0]) <= 1e-14 * dxt / rm[0]
    o.set("qt", np.zeros(shape))
    qr = np.zeros(shape, order="F")
    qr[1] = 1.0
    o.set("qr", qr)
    rc = o.vec("rc")
    # udx2m(k) = dxr / g2rm(k) with g2rm(k) = (rc(k+1) - rc(k)) dxr: 1/(rc(k+1)-rc(k))
    want = max(1.0 / ((rc[1] - rc[0]) * rm[0]), 1.0 / ((rc[2] - rc[1]) * rm[1]))
    assert abs(o.max_div() - want) <= 1e-12 * want
    o.close()


def test_velocity_stats_closed_forms(tc, oracle):
    """CalcVelocityStats.F90:29-60.  Uniform qt = U, qz = W and qr = V rc
    (so that q_r / r = V) give mean squares U^2, V^2, W^2 exactly, because the
    pvol(k) = rc(k+1)^2 - rc(k)^2 weights telescope to rext^2 - rint^2; vmax
    is the largest magnitude over kc = 1..nrm."""
    p, o = _reduction_oracle(tc, oracle)
    shape = o.shape
    U, V, W = 1.3, -0.4, 0.7
    rc = o.vec("rc")
    o.set("qt", np.full(shape, U))
    o.set("qr", np.broadcast_to(V * rc[: o.nr][:, None, None], shape))
    o.set("qz", np.full(shape, W))
    vmax, rms = o.velocity_stats()
    assert np.allclose(rms, [U * U, V * V, W * W], rtol=1e-13, atol=0)
    assert vmax[0] == U and vmax[2] == W
    assert vmax[1] == abs(V) * max(abs(rc[: o.nrm]))
    o.close()
