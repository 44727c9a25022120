"""Smooth analytic velocity fields and the continuous operators the explicit
and implicit terms discretise, for convergence checks of the oracle.

Coordinates: r in [rint, rext], theta periodic with 2 pi / lamb, z periodic
with alx3.  Variables as in the solver: v_theta, q_r = r v_r, v_z.
Derivatives of the analytic functions are taken with 4th-order central
differences (truncation ~1e-12, far below the O(h^2) grid errors checked).
"""
import numpy as np


class Field:
    def __init__(self, p):
        self.ri, self.ro = p["rint"], p["rext"]
        self.lam = p["lamb"]
        self.kz = 2 * np.pi / p["alx3"]

    def s(self, r):
        return (r - self.ri) / (self.ro - self.ri)

    def vt(self, r, t, z):
        return (0.3 * r + 0.2 / r + 0.05 * np.sin(np.pi * self.s(r)) *
                (np.cos(self.lam * t) + 0.5 * np.sin(self.kz * z) + 0.3 * np.sin(self.lam * t + self.kz * z)))

    def vr(self, r, t, z):
        return 0.04 * np.sin(np.pi * self.s(r)) ** 2 * np.cos(self.lam * t + self.kz * z + 0.3)

    def qr(self, r, t, z):
        return r * self.vr(r, t, z)

    def vz(self, r, t, z):
        return 0.03 * np.sin(np.pi * self.s(r)) * np.sin(self.lam * t + 0.2) * np.cos(self.kz * z)


def d1(f, x, h):
    return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)


def d2(f, x, h):
    return (-f(x + 2 * h) + 16 * f(x + h) - 30 * f(x) + 16 * f(x - h) - f(x - 2 * h)) / (12 * h * h)


H1, H2 = 1e-4, 1e-3


def cont_qt(F, ren, romeg, r, t, z):
    """ExplicitTermsQT: -[ (1/r) d(vt^2)/dt + (1/r) d(qr vt)/dr + qr vt/r^2 + d(vz vt)/dz ]
    + (1/Re)[ (1/r^2) d2vt/dt2 + d2vt/dz2 + (2/r^2) dvr/dt ] - romeg vr"""
    adv = (d1(lambda tt: F.vt(r, tt, z) ** 2, t, H1) / r
           + d1(lambda rr: F.qr(rr, t, z) * F.vt(rr, t, z), r, H1) / r
           + F.qr(r, t, z) * F.vt(r, t, z) / r ** 2
           + d1(lambda zz: F.vz(r, t, zz) * F.vt(r, t, zz), z, H1))
    vis = (d2(lambda tt: F.vt(r, tt, z), t, H2) / r ** 2 + d2(lambda zz: F.vt(r, t, zz), z, H2)
           + 2 * d1(lambda tt: F.vr(r, tt, z), t, H1) / r ** 2) / ren
    return -adv + vis - romeg * F.vr(r, t, z)


def cont_qr(F, ren, romeg, r, t, z):
    """ExplicitTermsQR: -[ (1/r) d(vt qr)/dt + d(qr^2/r)/dr + d(vz qr)/dz ] + vt^2
    + (1/Re)[ (1/r^2) d2qr/dt2 + d2qr/dz2 - (2/r) dvt/dt ] + romeg r vt"""
    adv = (d1(lambda tt: F.vt(r, tt, z) * F.qr(r, tt, z), t, H1) / r
           + d1(lambda rr: F.qr(rr, t, z) ** 2 / rr, r, H1)
           + d1(lambda zz: F.vz(r, t, zz) * F.qr(r, t, zz), z, H1))
    vis = (d2(lambda tt: F.qr(r, tt, z), t, H2) / r ** 2 + d2(lambda zz: F.qr(r, t, zz), z, H2)
           - 2 * d1(lambda tt: F.vt(r, tt, z), t, H1) / r) / ren
    return -adv + F.vt(r, t, z) ** 2 + vis + romeg * r * F.vt(r, t, z)


def cont_qz(F, ren, r, t, z):
    """ExplicitTermsQZ: -[ (1/r) d(vt vz)/dt + (1/r) d(qr vz)/dr + d(vz^2)/dz ]
    + (1/Re)[ (1/r^2) d2vz/dt2 + d2vz/dz2 ]"""
    adv = (d1(lambda tt: F.vt(r, tt, z) * F.vz(r, tt, z), t, H1) / r
           + d1(lambda rr: F.qr(rr, t, z) * F.vz(rr, t, z), r, H1) / r
           + d1(lambda zz: F.vz(r, t, zz) ** 2, z, H1))
    vis = (d2(lambda tt: F.vz(r, tt, z), t, H2) / r ** 2 + d2(lambda zz: F.vz(r, t, zz), z, H2)) / ren
    return -adv + vis


def staggered(o):
    """Node coordinates of the staggered grid (CreateGrid.F90:57-73,136-138)."""
    fs = o.vec("fs")
    dxt, dxz = fs[12], fs[13]
    rc, rm = o.vec("rc"), o.vec("rm")
    tc = np.arange(o.nthm) / dxt
    tm = (np.arange(o.nthm) + 0.5) / dxt
    zc = np.arange(o.nzm) / dxz
    zm = (np.arange(o.nzm) + 0.5) / dxz
    return rc, rm[: o.nrm], tc, tm, zc, zm


def sample_state(o, F):
    """qt at (rm, zm, tc), qr at (rc, zm, tm), qz at (rm, zc, tm)."""
    rc, rm, tc, tm, zc, zm = staggered(o)
    shape = o.shape
    qt = np.zeros(shape, order="F")
    qr = np.zeros(shape, order="F")
    qz = np.zeros(shape, order="F")
    R, Z, T = np.meshgrid(rm, zm, tc, indexing="ij")
    qt[: o.nrm] = F.vt(R, T, Z)
    R, Z, T = np.meshgrid(rc, zm, tm, indexing="ij")
    qr[:] = F.qr(R, T, Z)
    qr[0] = 0.0
    qr[-1] = 0.0
    R, Z, T = np.meshgrid(rm, zc, tm, indexing="ij")
    qz[: o.nrm] = F.vz(R, T, Z)
    return qt, qr, qz
