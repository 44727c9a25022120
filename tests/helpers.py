"""Shared set-up for the parity tests: identical states on the device and in
the oracle, and the comparison metrics."""
import numpy as np


def small_params(tc, nthm=32, nzm=24, nrm=33, **kw):
    p = tc.configs.small_64()
    p.update(nthm=nthm, nzm=nzm, nrm=nrm)
    p.update(kw)
    return p


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def random_state(p, seed, amp=1e-2, pressure=True, history=True):
    """Couette profile plus noise on every field the hot path reads."""
    rng = np.random.default_rng(seed)
    nr, nzm, nthm, nrm = p["nrm"] + 1, p["nzm"], p["nthm"], p["nrm"]
    shape = (nr, nzm, nthm)
    st = {}
    rint, rext = p["rint"], p["rext"]
    eta = rint / rext
    vi, ve = p["virotint"], p["virotext"]
    A = (ve - eta * eta * vi) / (1 - eta * eta)
    B = (vi - ve) * rint * rint / (1 - eta * eta)
    r = np.linspace(rint, rext, nr)
    rm = 0.5 * (r[1:] + r[:-1])
    qt = np.zeros(shape, order="F")
    qt[:nrm] = (A * rm + B / rm)[:, None, None] + amp * rng.uniform(-1, 1, (nrm, nzm, nthm))
    qr = np.zeros(shape, order="F")
    qr[1:nrm] = amp * rng.uniform(-1, 1, (nrm - 1, nzm, nthm))
    qz = np.zeros(shape, order="F")
    qz[:nrm] = amp * rng.uniform(-1, 1, (nrm, nzm, nthm))
    st.update(qt=qt, qr=qr, qz=qz)
    for name in ["pr", "dph", "ru1", "ru2", "ru3", "dq", "qcap"]:
        a = np.zeros(shape, order="F")
        if (name == "pr" and pressure) or (name.startswith("ru") and history) or name in ("dph", "dq", "qcap"):
            a[:nrm] = amp * rng.uniform(-1, 1, (nrm, nzm, nthm))
        st[name] = a
    return st


def load_both(dev, orc, st):
    for k, v in st.items():
        dev.set(k, v)
        orc.set(k, v)


def interior(a, nrm, kind="cell"):
    """The planes the reference defines for a field: 1..nrm for cell fields,
    1..nr for qr."""
    return a if kind == "qr" else a[:nrm]


def compare_fields(dev, orc, nrm, tol=1e-10):
    """Relative L2 of qt, qr, qz and the full pressure between a device
    solver and the oracle; asserts each is within tol.  Both fix the gauge
    of the singular (0,0) Poisson mode the same way (phi(nrm) = 0 on that
    mode), so the pressure is compared whole, mean profile included."""
    out = {}
    for n, sl in (("qt", slice(0, nrm)), ("qr", slice(0, nrm + 1)), ("qz", slice(0, nrm)),
                  ("pr", slice(0, nrm))):
        out[n] = rel_l2(dev.get(n)[sl], orc.get(n)[sl])
    for n, e in out.items():
        assert e <= tol, f"{n}: rel L2 {e:.3e} ({out})"
    return out


def start_pair(tc, oracle, p, seed, amp):
    """The seeded projected start on the device, copied into a fresh oracle."""
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    del qt, qr, qz
    dev.project()
    orc = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        orc.set(n, dev.get(n))
    return dev, orc


def write_bou(path, p):
    """A bou.in with the eleven header/value line pairs of ReadInputFile.F90:17-37."""
    g = lambda k: p.get(k, 0)
    lines = [
        "NTM NZM NRM NSST IDTV NREAD",
        f"{g('nthm')} {g('nzm')} {g('nrm')} {g('nsst')} {g('idtv')} {g('nread')}",
        "NTST TOUT TMAX IRESET WALLTIMEMAX",
        f"{g('ntst')} {g('tout')!r} {g('tmax')!r} {g('ireset')} {g('walltimemax')!r}",
        "ALX3D RINT REXT ISTR STR3 LAMB",
        f"{g('alx3')!r} {g('rint')!r} {g('rext')!r} {g('istr')} {g('str')!r} {g('lamb')!r}",
        "VIROTINT VIROTEXT AMPL_PERT FREQ_PERT",
        f"{g('virotint')!r} {g('virotext')!r} {g('sizepert')!r} {g('omegapert')!r}",
        "1/NU ROTP RESID CFLMAX",
        f"{g('ren')!r} {g('romeg')!r} {g('resid')!r} {g('limitcfl')!r}",
        "TSTA STAREAD",
        f"{g('tsta')!r} {g('starea')}",
        "DTMIN DTMAX VMAX CFLLIM",
        f"{g('dtmin')!r} {g('dtmax')!r} {g('limitvel')!r} {g('cfllim')!r}",
        "NOUT_RCUT OUT_THCUT OUT_STWAVE TFRAME2D",
        f"{g('nslabr')} {g('dpthcut')} {g('dpstave')} {g('tframe2d')!r}",
        "OUT_3DFIELD TFRAME3D",
        f"{g('dpfullf')} {g('tframe3d')!r}",
        "MINSTRT MINSTRZ MULTMIN",
        f"{g('minstrt')} {g('minstrz')} {g('multmin')}",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
