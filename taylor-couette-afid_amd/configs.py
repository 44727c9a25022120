"""bou.in parameter sets of the benchmark workloads.

Each dict carries the bou.in keys of ReadInputFile.F90:17-37 (lower case).
Values the workload descriptions leave open are fixed here: STR3 = 1.5
(ISTR = 4 tanh clustering at both walls), RESID = 1e-6, VMAX = 1e3, and for
the variable-dt runs DTMIN = 1e-4, DTMAX = 1e-2.  Fixed-dt runs set
IDTV = 0 and DTMIN = DTMAX = dt.
"""
import math

BASE = dict(
    nsst=3, nread=0, ntst=0, tout=1.0e30, tmax=1.0e30, ireset=1, walltimemax=1.0e30,
    istr=4, str=1.5, romeg=0.0, resid=1.0e-6, tsta=1.0e30, starea=0,
    limitvel=1.0e3, cfllim=1.0e30, nslabr=0, dpthcut=0, dpstave=0, tframe2d=1.0e30,
    dpfullf=0, tframe3d=1.0e30, minstrt=0, minstrz=0, multmin=0,
    sizepert=0.0, omegapert=0.0, virotext=0.0,
)


def small_64():
    """64x64x65, eta ~ 0.714, ALX3D = 2, LAMB = 4, 1/NU = 2500, fixed dt 1e-3, 200 steps."""
    p = dict(BASE)
    p.update(nthm=64, nzm=64, nrm=65, idtv=0, ntst=199, alx3=2.0, rint=1.0, rext=1.4, lamb=4.0,
             virotint=1.0, virotext=0.0, ren=2500.0, limitcfl=1.0e30, dtmin=1.0e-3, dtmax=1.0e-3)
    return p


def production_512():
    """512x512x385, eta ~ 0.714, ALX3D = 2 pi, LAMB = 20, 1/NU = 1e4, inner
    cylinder 1 + 0.1 sin(2 pi 0.05 t), RK3 variable dt at CFLMAX = 1.2,
    torque and divergence logged every step, 500 steps."""
    p = dict(BASE)
    p.update(nthm=512, nzm=512, nrm=385, idtv=1, ntst=499, alx3=2.0 * math.pi, rint=1.0,
             rext=1.4, lamb=20.0, virotint=1.0, virotext=0.0, sizepert=0.1, omegapert=0.05,
             ren=1.0e4, limitcfl=1.2, dtmin=1.0e-4, dtmax=1.0e-2, tout=1.0e-12)
    return p


def large_1024():
    """1024x1024x769, eta = 0.909, ALX3D = 2 pi, LAMB = 40, 1/NU = 3e4,
    VIROTEXT = -0.4 VIROTINT, RK3 variable dt, 100 steps."""
    p = dict(BASE)
    p.update(nthm=1024, nzm=1024, nrm=769, idtv=1, ntst=99, alx3=2.0 * math.pi, rint=1.0,
             rext=1.0 / 0.909, lamb=40.0, virotint=1.0, virotext=-0.4, ren=3.0e4,
             limitcfl=1.2, dtmin=1.0e-4, dtmax=1.0e-2)
    return p


def scaled(p, nthm, nzm, nrm):
    q = dict(p)
    q.update(nthm=nthm, nzm=nzm, nrm=nrm)
    return q


WORKLOADS = {"small_64": small_64, "production_512": production_512, "large_1024": large_1024}
SEEDS = {"small_64": 1001, "production_512": 2002, "large_1024": 3003}
AMPLITUDE = {"small_64": 1e-2, "production_512": 1e-2, "large_1024": 1e-3}
