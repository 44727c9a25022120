"""Pins of the oracle's restart onto another grid (interp.F90 as
ReadFlowField.F90 uses it), CPU only:
  - fields affine in theta and z, constant in r, are reproduced exactly at
    the new points (bilinear weights, bracket search, linear extrapolation
    of the halos), for each of the three staggerings interp.F90 assigns;
  - a field that depends on the old radial plane only is carried to the
    new plane as the value of the lower bracketing old plane (the
    reference's eight corner values all sit on that plane), checked with an
    independent searchsorted bracket;
  - more than twice the old resolution, and radial grids interp.F90's
    gridnew cannot build, are rejected.
"""
import numpy as np
import pytest


def params(tc, nthm, nzm, nrm, **kw):
    p = tc.configs.small_64()
    p.update(nthm=nthm, nzm=nzm, nrm=nrm, rint=1.0, rext=1.5, lamb=4.0, alx3=1.0, istr=0, str=0.0)
    p.update(kw)
    return p


def coords(n, lo, hi):
    """interp.F90 gridnew with ISTR 0: faces lo + (hi-lo)(i-1)/(n-1) for an
    origin at 0, midpoints between; index 0 and n+1 extrapolated."""
    c = np.zeros(n + 2)
    c[1:n + 1] = hi * np.arange(n) / (n - 1) + lo
    m = np.zeros(n + 2)
    m[1:n] = 0.5 * (c[1:n] + c[2:n + 1])
    m[n] = 2 * c[n] - m[n - 1]
    for v in (c, m):
        v[0] = 2 * v[1] - v[2]
        v[n + 1] = 2 * v[n] - v[n - 1]
    return c, m


# interp.F90:75-112: (theta, z, r) coordinate kinds of each variable
STAGGER = {"qt": ("c", "m", "m"), "qr": ("m", "c", "m"), "qz": ("m", "m", "c")}


@pytest.mark.parametrize("name", ["qt", "qr", "qz"])
@pytest.mark.parametrize("old_new", [((8, 6, 9), (16, 12, 17)), ((12, 10, 9), (16, 12, 13))])
def test_affine_in_theta_and_z_is_exact(tc, oracle, name, old_new):
    (tho, zo, ro), (thn, zn, rn) = old_new
    p = params(tc, thn, zn, rn)
    o = oracle.Oracle(p)
    thmax = 2 * np.pi / p["lamb"]
    a, b, c = 0.3, 0.7, -1.1
    kt, kz, _ = STAGGER[name]

    def at(nth, nz):
        tcv, tmv = coords(nth, 0.0, thmax)
        zcv, zmv = coords(nz, 0.0, p["alx3"])
        t = (tcv if kt == "c" else tmv)[1:nth]
        z = (zcv if kz == "c" else zmv)[1:nz]
        return a + b * t[None, :] + c * z[:, None]

    old2 = at(tho + 1, zo + 1)                   # (nzo-1, ntho-1)
    old = np.repeat(old2[None, :, :], ro, axis=0)
    got = o.regrid(name, old, 0, 0.0)
    want = at(thn + 1, zn + 1)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", ["qt", "qr", "qz"])
def test_radial_takes_the_lower_plane(tc, oracle, name):
    ro, rn = 9, 15
    p = params(tc, 8, 6, rn - 1, istr=6, str=2.0)
    o = oracle.Oracle(p)
    g = np.linspace(1.0, 2.0, ro) ** 2             # value of old plane k = 1..ro
    old = np.repeat(np.repeat(g[:, None, None], 5, axis=1), 7, axis=2)
    got = o.regrid(name, old, 0, 0.0)
    # independent brackets: old radial coordinates from ISTR 0 (rext*x+rint),
    # new ones from ISTR 6; lower plane = (first old >= new) - 1
    kr = STAGGER[name][2]
    rc0, rm0 = coords(ro, p["rint"], p["rext"])
    zold = (rc0 if kr == "c" else rm0)[1:ro + 1]
    rcn = np.zeros(rn + 2)
    nclip, nrmo = 2, rn + 4
    etazm = np.cos(np.pi * (np.arange(1, nrmo + 1) - 0.5) / nrmo)
    eta = etazm[nclip:nclip + rn]
    eta = eta / (0.5 * (eta[0] - eta[-1]))
    rcn[1:rn + 1] = (p["rext"] - p["rint"]) * (1 - eta) * 0.5 + p["rint"]
    rcn[1], rcn[rn] = p["rint"], p["rext"]
    rmn = np.zeros(rn + 2)
    rmn[1:rn] = 0.5 * (rcn[1:rn] + rcn[2:rn + 1])
    rmn[rn] = 2 * rcn[rn] - rmn[rn - 1]
    znew = (rcn if kr == "c" else rmn)[1:rn + 1]
    gext = np.concatenate([[2 * g[0] - g[1]], g])  # index 0 = extrapolated halo
    l = np.searchsorted(zold, znew, side="left") + 1
    lower = np.where(l > ro, ro - 1, l - 1)
    want = gext[lower]
    np.testing.assert_allclose(got[:, 0, 0], want, rtol=1e-14)
    assert np.allclose(got, got[:, :1, :1])


def test_rejections(tc, oracle):
    o = oracle.Oracle(params(tc, 16, 12, 17))
    with pytest.raises(ValueError):
        o.regrid("qt", np.zeros((9, 5, 7)), 0, 0.0)      # new nth 17 > 2 * old nth 8
    with pytest.raises(ValueError):
        o.regrid("qt", np.zeros((9, 11, 15)), 4, 1.5)    # ISTR 4: gridnew builds no grid
