"""Pins of the oracle's semi-implicit radial steps (CPU only).

ImplicitAndUpdateQT/QR/QZ + SolveImpEqnUpdate_QT/QR/QZ advance each velocity
component by dq, where, per radial line,

    rhs = (ga H + ro ru + (al/Re) L_e q - al G p) dt          (right-hand side)
    (I - beta al L_i) dq = rhs                                 (Crank-Nicolson)

H is the explicit term (dq, dph or qcap), ru the previous substep's H, G the
pressure gradient and L_e, L_i radial second-order operators (the explicit
and implicit one differ for q_r: A2e and A2, CreateGrid.F90:244-276).

Nothing below is read from the oracle's coefficient vectors.  Every radial
stencil is derived here from the operator it discretises, written in
conservative (flux) form on the staggered grid, through what that form does
to the functions 1, r and r^2: the fluxes telescope, so the images are closed
forms of the node positions.  For example A2 (the q_r = r v_r operator
d2/dr2 - (1/r) d/dr) maps 1 and r^2 to zero exactly, while A2e (the same
operator written on v_r = q_r / r) maps 1 to 1/(rc(k-1) rc(k+1)) - 1/rc(k)^2.
Three images fix a three-point stencil.  The pressure gradient is taken from
CorrectVelocity, which the Poisson tests pin (D(G(Solve)) = I).

The oracle is then run on random states, random masks and both time schemes,
and each updated line must satisfy the system above to rounding.  Each test
also checks that the plausible mistakes (A2 and A2e swapped, no-slip rows
where the masks ask for free slip, a consistent row scaling where the
reference has none, ga and ro swapped) would miss the bar by orders of
magnitude.

Reference behaviour the expectations keep (each cited where it is used):
  - the free-slip theta row keeps the no-slip coupling ap1j(1) to the second
    plane next to its stress-free diagonal (ImplicitAndUpdateQT.F90:47-52);
  - the right-hand sides of QT and QZ are scaled with the no-slip diagonal at
    the wall rows while the matrix rows use the blended one
    (SolveImpEqnUpdate_QT.F90:36-37,62-63; SolveImpEqnUpdate_QZ.F90:35-36,59-61);
  - the inner-wall acceleration term of QT uses the no-slip am1j(1) whatever
    the mask (SolveImpEqnUpdate_QT.F90:64);
  - the outer free-slip theta coefficients are never set (CreateGrid.F90:
    218-220, 236-238), so that row keeps only am1j(nrm);
  - the QZ history at row nrm is the new qcap (ImplicitAndUpdateQZ.F90:92-113).
"""
import numpy as np
import pytest

RK3 = {1: (8 / 15, 0.0), 2: (5 / 12, -17 / 60), 3: (3 / 4, -5 / 12)}
AB2 = (1.5, -0.5)


def base_params(tc, nthm=8, nzm=6, nrm=11, **kw):
    p = tc.configs.small_64()
    p.update(nthm=nthm, nzm=nzm, nrm=nrm, ren=10.0, romeg=0.0, rint=1.0, rext=1.5,
             lamb=4.0, alx3=1.0, istr=4, str=1.5, virotint=0.8, virotext=-0.3,
             sizepert=0.2, omegapert=0.7)
    p.update(kw)
    return p


def geometry(o):
    rc = o.vec("rc")[: o.nr]
    rm = 0.5 * (rc[1:] + rc[:-1])
    np.testing.assert_allclose(o.vec("rm")[: o.nrm], rm, rtol=1e-15, atol=0)
    return rc, rm


def stencil3(xm, x0, xp, t1, tr, tr2):
    """(am, ac, ap) of a three-point row from its images of 1, r, r^2,
    solved in the basis 1, r - x0, (r - x0)^2 for conditioning."""
    dm, dp = xm - x0, xp - x0
    n = len(x0)
    M = np.zeros((n, 3, 3))
    M[:, 0] = 1.0
    M[:, 1] = np.stack([dm, np.zeros(n), dp], axis=1)
    M[:, 2] = np.stack([dm * dm, np.zeros(n), dp * dp], axis=1)
    b = np.stack([t1, tr - x0 * t1, tr2 - 2 * x0 * tr + x0 * x0 * t1], axis=1)
    s = np.linalg.solve(M, b[..., None])[..., 0]
    return s[:, 0], s[:, 1], s[:, 2]


def flux_images(rc, xm, x0, xp, kc):
    """Images of 1, r, r^2 under (1/(x0 h)) [rc(k+1) df/dr|+ - rc(k) df/dr|-]
    with one-sided differences between nodes: the cylindrical radial
    Laplacian (1/r) d/dr (r d/dr) of a cell-centred field at plane kc."""
    h = rc[kc + 1] - rc[kc]
    t1 = np.zeros_like(x0)
    tr = (rc[kc + 1] - rc[kc]) / (x0 * h)
    tr2 = (rc[kc + 1] * (xp + x0) - rc[kc] * (x0 + xm)) / (x0 * h)
    return t1, tr, tr2


def cell_nodes(rc, rm):
    """neighbour nodes of the cell-centred planes 1..nrm: rm(k-1), rm(k+1),
    and the wall positions rc(1), rc(nr) beyond the first and last plane."""
    nrm = len(rm)
    xm = np.concatenate([[rc[0]], rm[:-1]])
    xp = np.concatenate([rm[1:], [rc[nrm]]])
    return xm, rm.copy(), xp


def op_a1(rc, rm):
    """A1 = (1/r) d/dr (r d/dr) - 1/r^2 for v_theta at rm; the wall rows
    take the wall as their outer node (no slip)."""
    xm, x0, xp = cell_nodes(rc, rm)
    kc = np.arange(len(rm))
    t1, tr, tr2 = flux_images(rc, xm, x0, xp, kc)
    return stencil3(xm, x0, xp, t1 - 1 / x0 ** 2, tr - 1 / x0, tr2 - 1.0)


def op_a3(rc, rm):
    """A3 = (1/r) d/dr (r d/dr) for v_z at rm; no-slip wall nodes."""
    xm, x0, xp = cell_nodes(rc, rm)
    kc = np.arange(len(rm))
    return stencil3(xm, x0, xp, *flux_images(rc, xm, x0, xp, kc))


def op_a2(rc, rm):
    """A2 on q_r at rc(k), k = 2..nrm: (rc/H) [dq/dr|+ / rm(k) - dq/dr|- / rm(k-1)]
    = d2q/dr2 - (1/r) dq/dr.  Images: 1 -> 0, r -> -rc/(rm(k) rm(k-1)),
    r^2 -> 0."""
    xm, x0, xp = rc[:-2], rc[1:-1], rc[2:]
    t1 = np.zeros_like(x0)
    tr = -x0 / (rm[1:] * rm[:-1])
    return stencil3(xm, x0, xp, t1, tr, np.zeros_like(x0))


def op_a2e(rc, rm):
    """A2e on q_r = r v_r at rc(k): r [(1/r) d/dr (r dv/dr) - v/r^2] with
    v = q/r differenced in flux form.  Images: 1 -> 1/(rc(k-1) rc(k+1))
    - 1/rc(k)^2, r -> -1/rc(k), r^2 -> 0."""
    xm, x0, xp = rc[:-2], rc[1:-1], rc[2:]
    t1 = 1.0 / (xm * xp) - 1.0 / (x0 * x0)
    tr = -1.0 / x0
    return stencil3(xm, x0, xp, t1, tr, np.zeros_like(x0))


def stress_free_theta_row(rc, rm):
    """Inner-wall row of A1 = (1/r^2) d/dr (r^3 d(v/r)/dr) with zero stress
    r^3 d(v/r)/dr = 0 at the wall:
    (1/(rm1^2 h)) rc(2)^3 [(v/r)(rm2) - (v/r)(rm1)] / (rm2 - rm1).
    Solid-body rotation v = r is its null mode, so ac rm1 + ap rm2 = 0, and
    ap = rc2^3 / (rm1^2 h rm2 (rm2 - rm1)); returns (ac, ap)."""
    h = rc[1] - rc[0]
    ap = rc[1] ** 3 / (rm[0] ** 2 * h * rm[1] * (rm[1] - rm[0]))
    return -ap * rm[1] / rm[0], ap


def free_slip_z_rows(rc, rm):
    """Rows of A3 with zero flux through the wall: (ac, ap) at the inner
    wall and (am, ac) at the outer wall, each mapping 1 to 0."""
    nrm = len(rm)
    h1 = rc[1] - rc[0]
    ap = rc[1] / (rm[0] * h1 * (rm[1] - rm[0]))
    hn = rc[nrm] - rc[nrm - 1]
    am = rc[nrm - 1] / (rm[nrm - 1] * hn * (rm[nrm - 1] - rm[nrm - 2]))
    return (-ap, ap), (am, -am)


def tridiag_apply(am, ac, ap, x):
    """(am x(k-1) + ac x(k) + ap x(k+1)) along axis 0, x zero beyond the ends;
    coefficient arrays broadcast against x."""
    y = ac * x
    y[1:] += (am * np.ones_like(x))[1:] * x[:-1]
    y[:-1] += (ap * np.ones_like(x))[:-1] * x[1:]
    return y


def col(v):
    return np.asarray(v)[:, None, None]


def random_state(o, seed):
    rng = np.random.default_rng(seed)
    st = {}
    for n in ("qt", "qr", "qz", "pr", "dq", "dph", "qcap", "ru1", "ru2", "ru3"):
        st[n] = np.asfortranarray(rng.uniform(-1, 1, o.shape))
    return st


def random_masks(o, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 2, (o.nzm, o.nthm)).astype(np.int32),
            rng.integers(0, 2, (o.nzm, o.nthm)).astype(np.int32))


def advance(oracle, p, st, comp, ns, masks, time=0.3, dt=0.05):
    """One ImplicitAndUpdate call on a fresh oracle; returns (dq, scalars)."""
    o = oracle.Oracle(p)
    for n, a in st.items():
        o.set(n, a)
    o.set_time(time, dt)
    o.begin_step()
    if masks is not None:
        o.set_masks(*masks)
    o.set_substep(ns)
    o.call("implicit_and_update_" + comp)
    sc = o.scalars()
    new = o.get(comp)
    o.close()
    return new - st[comp], new, sc


def gradient_term(oracle, p, st, comp, ns, time=0.3, dt=0.05):
    """-al dt G pr as CorrectVelocity applies it (CorrectVelocity.F90:35-40)."""
    o = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        o.set(n, np.zeros(o.shape))
    o.set("dph", st["pr"])
    o.set_time(time, dt)
    o.set_substep(ns)
    o.call("correct_velocity")
    g = o.get(comp)
    o.close()
    return g


def scheme_coefficients(sc, nsst, ns):
    ga, ro = AB2 if nsst == 1 else RK3[ns]
    assert sc["ga"] == ga and sc["ro"] == ro and sc["al"] == ga + ro
    return ga, ro, ga + ro


CASES = [(3, 1), (3, 2), (3, 3), (1, 1)]  # (nsst, substep)


def rel_residual(lhs, rhs):
    return np.abs(lhs - rhs).max() / np.abs(rhs).max()


@pytest.mark.parametrize("nsst,ns", CASES)
@pytest.mark.parametrize("mask", ["noslip", "mixed"])
def test_theta_helmholtz(tc, oracle, nsst, ns, mask):
    """ImplicitAndUpdateQT.F90:40-118 + SolveImpEqnUpdate_QT.F90:34-78."""
    p = base_params(tc, nsst=nsst)
    o = oracle.Oracle(p)
    rc, rm = geometry(o)
    nrm = o.nrm
    st = random_state(o, 11 + ns)
    masks = None if mask == "noslip" else random_masks(o, 7)
    d, _, sc = advance(oracle, p, st, "qt", ns, masks)
    d = d[:nrm]
    ga, ro, al = scheme_coefficients(sc, nsst, ns)
    dt, bal = sc["dt"], sc["beta"] * sc["al"]
    assert sc["beta"] == dt / p["ren"] * 0.5
    q = st["qt"][:nrm]
    grad = gradient_term(oracle, p, st, "qt", ns)[:nrm]
    icyl, ocyl = masks if masks is not None else (np.ones((o.nzm, o.nthm)),) * 2
    fi, fo = np.asarray(icyl, float)[None], np.asarray(ocyl, float)[None]

    am, ac, ap = op_a1(rc, rm)
    ac_sf, _ = stress_free_theta_row(rc, rm)
    # explicit operator per line: the wall nodes carry the wall velocity
    w_in = p["virotint"] + sc["vinner_t"]
    w_out = p["virotext"]

    def explicit(am, ac, ap, ac_sf):
        Lq = tridiag_apply(col(am), col(ac), col(ap), q.copy())
        Lq[0] = ap[0] * q[1] + (ac[0] * fi + ac_sf * (1 - fi))[0] * q[0] + am[0] * w_in * fi[0]
        Lq[-1] = am[-1] * q[-2] + ac[-1] * fo[0] * q[-1] + ap[-1] * w_out * fo[0]
        return Lq

    def implicit_lhs(am, ac, ap, ac_sf, dd):
        acm = np.broadcast_to(col(ac), dd.shape).copy()
        acm[0] = ac[0] * fi[0] + ac_sf * (1 - fi[0])
        acm[-1] = ac[-1] * fo[0]
        y = (1 - bal * acm) * dd
        y[1:] -= bal * col(am)[1:] * dd[:-1]
        y[:-1] -= bal * col(ap)[:-1] * dd[1:]
        return y, acm

    def expected(am, ac, ap, ac_sf, quirk=True, g=ga, r=ro):
        rhs = (g * st["dq"][:nrm] + r * st["ru1"][:nrm]) * dt + 2 * bal * explicit(am, ac, ap, ac_sf) + grad
        rhs[0] += bal * am[0] * sc["dvinner_t"]
        lhs, acm = implicit_lhs(am, ac, ap, ac_sf, d)
        scale = (1 - bal * acm) / (1 - bal * col(ac)) if quirk else 1.0
        return lhs, rhs * scale

    lhs, rhs = expected(am, ac, ap, ac_sf)
    assert rel_residual(lhs, rhs) < 1e-13, rel_residual(lhs, rhs)
    # the stencils have power: each plausible mistake misses by orders of magnitude
    if ns > 1 or nsst == 1:
        assert rel_residual(*expected(am, ac, ap, ac_sf, g=ro, r=ga)) > 1e-3
    if mask == "mixed":
        assert rel_residual(*expected(am, ac, ap, ac[0])) > 1e-6  # no-slip row kept
        assert rel_residual(*expected(am, ac, ap, ac_sf, quirk=False)) > 1e-6
    am_e, ac_e, ap_e = op_a3(rc, rm)  # the Laplacian without -1/r^2
    assert rel_residual(*expected(am_e, ac_e, ap_e, ac_sf)) > 1e-6


@pytest.mark.parametrize("nsst,ns", CASES)
def test_radial_helmholtz_explicit_a2e_implicit_a2(tc, oracle, nsst, ns):
    """ImplicitAndUpdateQR.F90:39-71 + SolveImpEqnUpdate_QR.F90:25-67: the
    right-hand side applies A2e (wall planes qr(1), qr(nr) as stored), the
    solve inverts I - beta al A2 with identity rows at both walls, and the
    walls are reset to zero."""
    p = base_params(tc, nsst=nsst)
    o = oracle.Oracle(p)
    rc, rm = geometry(o)
    nr = o.nr
    st = random_state(o, 21 + ns)
    d, new, sc = advance(oracle, p, st, "qr", ns, None)
    assert np.all(new[0] == 0.0) and np.all(new[nr - 1] == 0.0)
    ga, ro, al = scheme_coefficients(sc, nsst, ns)
    dt, bal = sc["dt"], sc["beta"] * sc["al"]
    q = st["qr"]
    grad = gradient_term(oracle, p, st, "qr", ns)[1:nr - 1]
    dd = np.zeros_like(q)
    dd[1:nr - 1] = d[1:nr - 1]

    def expected(op_rhs, op_lhs, g=ga, r=ro):
        am, ac, ap = op_rhs(rc, rm)
        Lq = col(am) * q[:-2] + col(ac) * q[1:-1] + col(ap) * q[2:]
        rhs = (g * st["dph"][1:nr - 1] + r * st["ru2"][1:nr - 1]) * dt + 2 * bal * Lq + grad
        am, ac, ap = op_lhs(rc, rm)
        lhs = dd[1:-1] - bal * (col(am) * dd[:-2] + col(ac) * dd[1:-1] + col(ap) * dd[2:])
        return lhs, rhs

    lhs, rhs = expected(op_a2e, op_a2)
    assert rel_residual(lhs, rhs) < 1e-13, rel_residual(lhs, rhs)
    assert rel_residual(*expected(op_a2, op_a2)) > 1e-5
    assert rel_residual(*expected(op_a2e, op_a2e)) > 1e-5
    assert rel_residual(*expected(op_a2, op_a2e)) > 1e-5
    if ns > 1 or nsst == 1:
        assert rel_residual(*expected(op_a2e, op_a2, g=ro, r=ga)) > 1e-3


@pytest.mark.parametrize("nsst,ns", CASES)
@pytest.mark.parametrize("mask", ["noslip", "mixed"])
def test_axial_helmholtz(tc, oracle, nsst, ns, mask):
    """ImplicitAndUpdateQZ.F90:36-119 + SolveImpEqnUpdate_QZ.F90:32-68."""
    p = base_params(tc, nsst=nsst)
    o = oracle.Oracle(p)
    rc, rm = geometry(o)
    nrm = o.nrm
    st = random_state(o, 31 + ns)
    masks = None if mask == "noslip" else random_masks(o, 9)
    d, _, sc = advance(oracle, p, st, "qz", ns, masks)
    d = d[:nrm]
    ga, ro, al = scheme_coefficients(sc, nsst, ns)
    dt, bal = sc["dt"], sc["beta"] * sc["al"]
    q = st["qz"][:nrm]
    grad = gradient_term(oracle, p, st, "qz", ns)[:nrm]
    icyl, ocyl = masks if masks is not None else (np.ones((o.nzm, o.nthm)),) * 2
    fi, fo = np.asarray(icyl, float), np.asarray(ocyl, float)
    am, ac, ap = op_a3(rc, rm)
    (ac_fi, _), (_, ac_fo) = free_slip_z_rows(rc, rm)

    def wall_diag(ac, fsi, fso):
        acm = np.broadcast_to(col(ac), d.shape).copy()
        acm[0] = ac[0] * fi + fsi * (1 - fi)
        acm[-1] = ac[-1] * fo + fso * (1 - fo)
        return acm

    def expected(fsi, fso, quirk=True, wall_history=True, g=ga, r=ro):
        acm = wall_diag(ac, fsi, fso)
        Lq = col(am) * 0 + acm * q
        Lq[1:] += col(am)[1:] * q[:-1]
        Lq[:-1] += col(ap)[:-1] * q[1:]
        hist = r * st["ru3"][:nrm]
        if wall_history:
            hist[-1] = r * st["qcap"][nrm - 1]
        rhs = (g * st["qcap"][:nrm] + hist) * dt + 2 * bal * Lq + grad
        lhs = (1 - bal * acm) * d
        lhs[1:] -= bal * col(am)[1:] * d[:-1]
        lhs[:-1] -= bal * col(ap)[:-1] * d[1:]
        scale = (1 - bal * acm) / (1 - bal * col(ac)) if quirk else 1.0
        return lhs, rhs * scale

    lhs, rhs = expected(ac_fi, ac_fo)
    assert rel_residual(lhs, rhs) < 1e-13, rel_residual(lhs, rhs)
    if ns > 1 or nsst == 1:
        assert rel_residual(*expected(ac_fi, ac_fo, wall_history=False)) > 1e-6
        assert rel_residual(*expected(ac_fi, ac_fo, g=ro, r=ga)) > 1e-3
    if mask == "mixed":
        assert rel_residual(*expected(ac[0], ac[-1])) > 1e-6
        assert rel_residual(*expected(ac_fi, ac_fo, quirk=False)) > 1e-6


def test_free_slip_rows_have_the_wall_null_modes(tc, oracle):
    """A uniform q_z is an exact null mode of a free-slip wall row and not of
    a no-slip one (zero wall flux against a zero wall value).  Through the
    whole ImplicitAndUpdateQZ step: with free slip at both walls, uniform qz
    and nothing else, the right-hand side vanishes to rounding and qz stays
    uniform; with no slip it does not."""
    p = base_params(tc)
    out = {}
    for case in ("free", "noslip"):
        o = oracle.Oracle(p)
        qz = np.zeros(o.shape, order="F")
        qz[: o.nrm] = 0.6
        st = {n: np.zeros(o.shape, order="F") for n in ("qt", "qr", "pr", "qcap", "ru3")}
        st["qz"] = qz
        z = np.zeros((o.nzm, o.nthm), dtype=np.int32)
        masks = (z, z) if case == "free" else (z + 1, z + 1)
        d, _, _ = advance(oracle, p, st, "qz", 2, masks)
        out[case] = np.abs(d[: o.nrm]).max()
    assert out["free"] < 1e-14 * 0.6, out
    assert out["noslip"] > 1e-3, out


def test_rk3_and_ab2_coefficients(tc, oracle):
    """InitTimeMarchScheme.F90:16-48: RK3 gamma = (8/15, 5/12, 3/4),
    rho = (0, -17/60, -5/12); AB2 gamma = 3/2, rho = -1/2; alpha = gamma + rho
    (each alpha sums to the step: 8/15 + 2/15 + 1/3 = 1)."""
    for nsst, cases in ((3, (1, 2, 3)), (1, (1,))):
        o = oracle.Oracle(base_params(tc, nsst=nsst))
        tot = 0.0
        for ns in cases:
            o.set_substep(ns)
            sc = o.scalars()
            ga, ro, al = scheme_coefficients(sc, nsst, ns)
            tot += al
        assert abs(tot - 1.0) < 1e-15


def _mode_run(tc, oracle, comp, nsst, dt, tend, history=True):
    """A small-amplitude mode between no-slip walls that the projection and
    the nonlinear terms leave alone to rounding: v_theta(r, z) (no theta
    dependence, so no theta pressure gradient; its centrifugal forcing of q_r
    is of second order in the amplitude) or v_z(r, theta) (divergence free,
    no advection).  It decays under explicit theta/z diffusion and implicit
    radial diffusion.  With history=True the first step's RK history holds
    the explicit term of the starting field, as a started AB2 scheme would."""
    p = base_params(tc, nthm=16, nzm=16, nrm=17, nsst=nsst, idtv=0, dtmin=dt, dtmax=dt,
                    ren=20.0, virotint=0.0, virotext=0.0, sizepert=0.0, limitcfl=1e30,
                    ntst=10 ** 6, tout=1e30, istr=0, alx3=2.0)
    o = oracle.Oracle(p)
    rc, rm = geometry(o)
    amp = 1e-7
    prof = amp * np.sin(np.pi * (rm - 1.0) / 0.5)[:, None, None]
    q = np.zeros(o.shape, order="F")
    if comp == "qt":
        zm = (np.arange(o.nzm) + 0.5) / o.nzm * 2 * np.pi
        q[: o.nrm] = prof * np.cos(zm)[None, :, None] * np.ones((1, 1, o.nthm))
    else:
        th = (np.arange(o.nthm) + 0.5) / o.nthm * 2 * np.pi
        q[: o.nrm] = prof * np.cos(th)[None, None, :] * np.ones((1, o.nzm, 1))
    o.set(comp, q)
    if history:
        o.call("explicit_terms_" + comp)
        h, ru = ("dq", "ru1") if comp == "qt" else ("qcap", "ru3")
        o.set(ru, o.get(h))
    for n in range(int(round(tend / dt))):
        c, _ = o.step(n)
        assert c == 0
    out = o.get(comp)
    o.close()
    return out / amp


def _order_ratios(tc, oracle, comp, nsst, history=True, dts=(0.02, 0.01, 0.005, 0.0025)):
    sols = [_mode_run(tc, oracle, comp, nsst, dt, 0.4, history) for dt in dts]
    diffs = [np.abs(sols[i] - sols[i + 1]).max() for i in range(len(dts) - 1)]
    return tuple(diffs[i] / diffs[i + 1] for i in range(len(diffs) - 1))


@pytest.mark.parametrize("nsst", [3, 1])
def test_time_scheme_is_second_order(tc, oracle, nsst):
    """The IMEX pair (RK3 or AB2 explicit, Crank-Nicolson radial,
    InitTimeMarchScheme.F90:16-48 with TimeMarcher.F90:20-29) is second
    order in dt: successive differences shrink by 4 when dt halves.  A wrong
    gamma/rho split makes the explicit part first order (ratio 2) or
    inconsistent (ratio 1)."""
    r1, r2 = _order_ratios(tc, oracle, "qt", nsst)
    assert 3.6 < r1 < 4.4 and 3.8 < r2 < 4.2, (nsst, r1, r2)


def test_reference_first_order_parts(tc, oracle):
    """Two reference behaviours cost an order in dt, and the oracle keeps
    both: AB2 started with an empty history (the reference has no start-up
    step: its first step weights the explicit term 3/2) is first order, and
    so is v_z under RK3, because row nrm of ImplicitAndUpdateQZ takes
    (ga + ro) qcap in place of ga qcap + ro ru3 (ImplicitAndUpdateQZ.F90:
    92-113)."""
    r1, r2 = _order_ratios(tc, oracle, "qt", 1, history=False)
    assert 1.8 < r2 < 2.2, (r1, r2)
    fine = (0.005, 0.0025, 0.00125, 0.000625)
    r1, r2 = _order_ratios(tc, oracle, "qz", 3, dts=fine)
    assert r1 > r2 and r2 < 3.0, (r1, r2)
    r1, r2 = _order_ratios(tc, oracle, "qt", 3, dts=fine)
    assert 3.9 < r2 < 4.1, (r1, r2)
