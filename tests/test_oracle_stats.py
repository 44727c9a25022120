"""Pins of the oracle's statistics (CPU only): CalcStats with the vorticity
of CalcVorticity, and CalcDissipation.

Each check is a fact the mathematics fixes, chosen so that a dropped term,
a wrong sign, a wrong metric factor or an index off by one fails it:
  - laminar Couette flow has nute -> 1 (the normalisation lamdissip is its
    dissipation), converging as the radial grid is refined;
  - the dissipation functional of a smooth 3-D field with all six strain
    components nonzero converges to its continuous volume mean;
  - vorticity of solid-body rotation is 2*Omega on a uniform radial grid;
    vorticity of u_z = c r is -c; the discrete z- and theta-derivatives of
    sinusoids are the exact central-difference factors sin(a h)/h;
  - theta-means of theta-constant fields are the fields themselves, and the
    accumulators add up over calls (nstatsamples counts them);
  - vtvz_mean pairs qt with qz(k,j) + qr(k,j+1), the reference's own
    expression (StatRoutines.F90:59-60).
"""
import numpy as np
import pytest

import analytic as A


def params(tc, nthm, nzm, nrm, **kw):
    p = tc.configs.small_64()
    p.update(nthm=nthm, nzm=nzm, nrm=nrm, ren=10.0, romeg=0.0, rint=1.0, rext=1.5,
             lamb=4.0, alx3=1.0, istr=4, str=1.5, virotint=1.0, virotext=-0.3)
    p.update(kw)
    return p


def couette(p):
    ri, ro, vi, vo = p["rint"], p["rext"], p["virotint"], p["virotext"]
    al = (vo * ro - vi * ri) / (ro * ro - ri * ri)
    bl = (vi - al * ri) * ri
    return al, bl


def lamdissip(p):
    ri, ro, vi, vo = p["rint"], p["rext"], p["virotint"], p["virotext"]
    return 4 * ro ** 2 * ri ** 2 / (ro + ri) ** 2 * ((vi / ri - vo / ro) / (ro - ri)) ** 2


@pytest.mark.parametrize("istr", [0, 4])
def test_couette_dissipation_tends_to_one(tc, oracle, istr):
    errs = []
    for n in (16, 32, 64):
        p = params(tc, 4, 4, n, istr=istr)
        o = oracle.Oracle(p)
        al, bl = couette(p)
        rm = o.vec("rm")[:n]
        qt = np.zeros(o.shape, order="F")
        qt[:n] = (al * rm + bl / rm)[:, None, None]
        o.set("qt", qt)
        errs.append(abs(o.calc_dissipation() - 1.0))
    assert errs[0] < 0.05, errs
    assert errs[1] / errs[0] < 0.6 and errs[2] / errs[1] < 0.6, errs


def cont_dissipation_mean(p, F, nq=48):
    """Volume mean of 2(e11^2+e22^2+e33^2)+e12^2+e13^2+e23^2 of the analytic
    field, with the strain components of the cylindrical velocity gradient
    as CalcDissipation.F90 writes them."""
    ri, ro = p["rint"], p["rext"]
    x, w = np.polynomial.legendre.leggauss(nq)
    r = 0.5 * (ro - ri) * x + 0.5 * (ro + ri)
    wr = 0.5 * (ro - ri) * w
    th = np.arange(nq) * (2 * np.pi / p["lamb"]) / nq
    z = np.arange(nq) * p["alx3"] / nq
    R, T, Z = np.meshgrid(r, th, z, indexing="ij")
    h = 1e-4
    d1 = A.d1
    vt = lambda rr, tt, zz: F.vt(rr, tt, zz)
    e11 = d1(lambda tt: F.vt(R, tt, Z), T, h) / R + F.vr(R, T, Z) / R
    e22 = d1(lambda rr: F.vr(rr, T, Z), R, h)
    e33 = d1(lambda zz: F.vz(R, T, zz), Z, h)
    # the reference's e12 (CalcDissipation.F90:78-80) takes (1/r) d q_r/d theta
    # = d v_r/d theta, without the 1/r of the physical 2 e_r_theta
    e12 = R * d1(lambda rr: vt(rr, T, Z) / rr, R, h) + d1(lambda tt: F.vr(R, tt, Z), T, h)
    e13 = d1(lambda zz: F.vt(R, T, zz), Z, h) + d1(lambda tt: F.vz(R, tt, Z), T, h) / R
    e23 = d1(lambda rr: F.vz(rr, T, Z), R, h) + d1(lambda zz: F.vr(R, T, zz), Z, h)
    d = 2 * (e11 ** 2 + e22 ** 2 + e33 ** 2) + e12 ** 2 + e13 ** 2 + e23 ** 2
    num = np.einsum("k,kij->", wr * r, d) / (nq * nq)
    return num / (0.5 * (ro * ro - ri * ri))


def test_dissipation_converges_for_a_3d_field(tc, oracle):
    p0 = params(tc, 8, 8, 8)
    F = A.Field(p0)
    ref = cont_dissipation_mean(p0, F) / lamdissip(p0)
    errs = []
    for n in (16, 32, 64):
        p = params(tc, n, n, n)
        o = oracle.Oracle(p)
        qt, qr, qz = A.sample_state(o, F)
        o.set("qt", qt); o.set("qr", qr); o.set("qz", qz)
        errs.append(abs(o.calc_dissipation() - ref) / ref)
    assert errs[0] < 0.1, errs
    assert errs[1] / errs[0] < 0.6 and errs[2] / errs[1] < 0.6, errs


def test_vorticity_solid_body_and_linear_axial(tc, oracle):
    """uniform radial grid: v_theta = Omega r gives omega_z = 2 Omega in the
    interior; v_z = c r gives omega_theta = -c; omega_r = 0."""
    n = 24
    p = params(tc, 8, 6, n, istr=0)
    o = oracle.Oracle(p)
    om, c = 0.7, 0.3
    rm = o.vec("rm")[:n]
    qt = np.zeros(o.shape, order="F")
    qz = np.zeros(o.shape, order="F")
    qt[:n] = (om * rm)[:, None, None]
    qz[:n] = (c * rm)[:, None, None]
    o.set("qt", qt); o.set("qz", qz)
    o.calc_stats()
    st, ns = o.stats()
    assert ns == 1
    inner = slice(1, n - 1)
    np.testing.assert_allclose(st["wz_mean"][inner], 2 * om, rtol=1e-12)
    np.testing.assert_allclose(st["wt_mean"][inner], -c, rtol=1e-12)
    assert np.abs(st["wr_mean"]).max() < 1e-14
    np.testing.assert_allclose(st["wz_rms"][inner], (2 * om) ** 2, rtol=1e-12)


def test_vorticity_of_sinusoids(tc, oracle):
    """u_r = sin(a z) gives d u_r/dz = cos(a z) sin(a h)/h exactly at the
    cell centres; u_z = sin(m theta) gives (1/r) d u_z/d theta whose
    theta-mean square is (sin(m dt)/(dt rm))^2 / 2."""
    nthm, nzm, n = 16, 12, 10
    p = params(tc, nthm, nzm, n)
    o = oracle.Oracle(p)
    rc, rm, tc_, tm, zc, zm = A.staggered(o)
    a = 2 * np.pi / p["alx3"]
    hz = p["alx3"] / nzm
    qr = np.zeros(o.shape, order="F")
    qr[:] = rc[:, None, None] * np.sin(a * zm)[None, :, None]
    o.set("qr", qr)
    o.calc_stats()
    st, _ = o.stats()
    want = np.cos(a * zm) * np.sin(a * hz) / hz
    np.testing.assert_allclose(st["wt_mean"], np.broadcast_to(want, (n, nzm)), atol=1e-12)

    o2 = oracle.Oracle(p)
    m = p["lamb"]
    dt = 2 * np.pi / p["lamb"] / nthm
    qz = np.zeros(o2.shape, order="F")
    qz[:n] = np.sin(m * tm)[None, None, :]
    o2.set("qz", qz)
    o2.calc_stats()
    st2, _ = o2.stats()
    want_rms = (np.sin(m * dt) / (dt * rm)) ** 2 / 2
    np.testing.assert_allclose(st2["wr_rms"], np.broadcast_to(want_rms[:, None], (n, nzm)),
                               rtol=1e-12)
    assert np.abs(st2["wr_mean"]).max() < 1e-13


def test_means_rms_and_accumulation(tc, oracle):
    n, nzm = 8, 6
    p = params(tc, 8, nzm, n)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(5)
    qt = np.zeros(o.shape, order="F")
    qr = np.zeros(o.shape, order="F")
    qz = np.zeros(o.shape, order="F")
    pr = np.zeros(o.shape, order="F")
    # theta-constant fields
    qt[:n] = rng.uniform(-1, 1, (n, nzm))[:, :, None]
    qr[1:n] = rng.uniform(-1, 1, (n - 1, nzm))[:, :, None]
    qz[:n] = rng.uniform(-1, 1, (n, nzm))[:, :, None]
    pr[:n] = rng.uniform(-1, 1, (n, nzm))[:, :, None]
    for name, f in (("qt", qt), ("qr", qr), ("qz", qz), ("pr", pr)):
        o.set(name, f)
    usrc = 1.0 / o.vec("rc")
    o.calc_stats()
    o.calc_stats()
    st, ns = o.stats()
    assert ns == 2
    t, r, z, pp = qt[:n, :, 0], qr[:, :, 0], qz[:n, :, 0], pr[:n, :, 0]
    vr = r * usrc[:, None]
    np.testing.assert_allclose(st["vt_mean"], 2 * t, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["vr_mean"], 2 * vr[:n], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["vz_mean"], 2 * z, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["pr_mean"], 2 * pp, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["vt_rms"], 2 * t * t, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["vr_rms"], 2 * vr[:n] ** 2, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(st["pr_rms"], 2 * pp * pp, rtol=1e-13, atol=1e-15)
    # face averages: radial faces k and kpv(k) (kpv(nrm) = nrm), z face j+1 (periodic)
    kp = np.minimum(np.arange(n) + 1, n - 1)
    vr_c = 0.5 * (vr[np.arange(n)] + vr[kp])
    np.testing.assert_allclose(st["vtvr_mean"], 2 * t * vr_c, rtol=1e-12, atol=1e-15)
    z_up = np.roll(z, -1, axis=1)
    np.testing.assert_allclose(st["vrvz_mean"], 2 * 0.5 * (z + z_up) * vr_c, rtol=1e-12, atol=1e-15)
    r_up = np.roll(r[:n], -1, axis=1)
    np.testing.assert_allclose(st["vtvz_mean"], 2 * t * 0.5 * (z + r_up), rtol=1e-12, atol=1e-15)


def test_step_runs_stats_after_tsta(tc, oracle):
    """main.F90:188-190: statistics only once time > TSTA, at logging steps"""
    p = params(tc, 8, 6, 8, tsta=0.015, dtmax=0.01, dtmin=0.01, idtv=0, tout=1e-9, ntst=4)
    o = oracle.Oracle(p)
    seen = []
    for n in range(4):
        o.step(n)
        seen.append(o.stats()[1])
    assert seen == [0, 1, 2, 3]
    assert np.isfinite(o.last_nute())


def test_slab_dumps_are_the_slices(tc, oracle):
    """DumpRCuts / DumpThCuts / DumpStreamwiseAvg against numpy slices and
    theta-means of the oracle's own fields (indices and face averages)."""
    n, nzm, nthm = 9, 6, 10
    p = params(tc, nthm, nzm, n)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(9)
    f = {}
    for name in ("qt", "qr", "qz"):
        a = np.zeros(o.shape, order="F")
        a[:] = rng.uniform(-1, 1, o.shape)
        o.set(name, a)
        f[name] = o.get(name)
    usrc = 1.0 / o.vec("rc")
    for k in (1, 4, n):
        vt, vr, vz = o.dump_rcut(k)
        assert np.array_equal(vt, f["qt"][k - 1]) and np.array_equal(vz, f["qz"][k - 1])
        want = 0.5 * (f["qr"][k - 1] * usrc[k - 1] + f["qr"][k] * usrc[k])
        np.testing.assert_allclose(vr, want, rtol=1e-15, atol=1e-15)
    vt, vr, vz = o.dump_thcut()
    assert np.array_equal(vt, f["qt"][:n, :, 0]) and np.array_equal(vz, f["qz"][:n, :, 0])
    np.testing.assert_allclose(vr, f["qr"][:n, :, 0] * usrc[:n, None], rtol=1e-15)
    vt, vr, vz = o.dump_streamwise_avg()
    np.testing.assert_allclose(vt, f["qt"][:n].mean(axis=2), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(vr, (f["qr"][:n] * usrc[:n, None, None]).mean(axis=2),
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(vz, f["qz"][:n].mean(axis=2), rtol=1e-13, atol=1e-15)
