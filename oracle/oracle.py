"""ctypes view of liboracle.so, the CPU restatement of the reference time step.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference result.  Fields are numpy arrays of shape
(nr, nzm, nthm) in Fortran order, i.e. the reference's (1:nr,1:nzm,1:nthm)
without halos.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_PATH = os.path.join(_HERE, "liboracle.so")

FIELDS = {"qt": 0, "qr": 1, "qz": 2, "pr": 3, "dph": 4, "rhs": 5, "dq": 6,
          "qcap": 7, "ru1": 8, "ru2": 9, "ru3": 10}
VECS = ["rc", "rm", "g2rc", "g2rm", "usrc", "usrm", "udx2rm", "udx2m", "udx2c",
        "rudx2c", "udx1vis1", "udx1vis2", "udx1qm", "ap1j", "am1j", "ac1j",
        "ap2j", "am2j", "ac2j", "ap2je", "am2je", "ac2je", "ap3j", "am3j",
        "ac3j", "apph", "amph", "acph", "ak1", "ak3", "fs", "pvol"]
# stat_arrays in WriteStats order (StatRoutines.F90:97-152), file stems
STATS = ["vt_mean", "vr_mean", "vz_mean", "wt_mean", "wr_mean", "wz_mean",
         "vt_rms", "vr_rms", "vz_rms", "wt_rms", "wr_rms", "wz_rms",
         "vtvr_mean", "vtvz_mean", "vrvz_mean", "pr_mean", "pr_rms", "disste"]

PARAM_FIELDS = [
    ("nthm", ctypes.c_int), ("nzm", ctypes.c_int), ("nrm", ctypes.c_int),
    ("nsst", ctypes.c_int), ("idtv", ctypes.c_int), ("nread", ctypes.c_int),
    ("ntst", ctypes.c_int), ("tout", ctypes.c_double), ("tmax", ctypes.c_double),
    ("ireset", ctypes.c_int), ("walltimemax", ctypes.c_double),
    ("alx3", ctypes.c_double), ("rint", ctypes.c_double), ("rext", ctypes.c_double),
    ("istr", ctypes.c_int), ("str", ctypes.c_double), ("lamb", ctypes.c_double),
    ("virotint", ctypes.c_double), ("virotext", ctypes.c_double),
    ("sizepert", ctypes.c_double), ("omegapert", ctypes.c_double),
    ("ren", ctypes.c_double), ("romeg", ctypes.c_double), ("resid", ctypes.c_double),
    ("limitcfl", ctypes.c_double), ("tsta", ctypes.c_double), ("starea", ctypes.c_int),
    ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double), ("limitvel", ctypes.c_double),
    ("cfllim", ctypes.c_double), ("nslabr", ctypes.c_int), ("dpthcut", ctypes.c_int),
    ("dpstave", ctypes.c_int), ("tframe2d", ctypes.c_double), ("dpfullf", ctypes.c_int),
    ("tframe3d", ctypes.c_double), ("minstrt", ctypes.c_int), ("minstrz", ctypes.c_int),
    ("multmin", ctypes.c_int),
]


class OrParams(ctypes.Structure):
    _fields_ = PARAM_FIELDS


def make_params(d):
    p = OrParams()
    for name, _ in PARAM_FIELDS:
        setattr(p, name, d.get(name, 0))
    return p


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def use_library(path):
    """Load the restatement from another build of the same source (the
    benchmark's -O3 -march=native build); call before any other use."""
    global _PATH
    if _LIB is not None:
        raise RuntimeError("the oracle library is already loaded")
    _PATH = path


def lib():
    global _LIB
    if _LIB is None:
        path = _PATH
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.POINTER(OrParams)]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_get_vec.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int]
        L.or_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.or_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.or_get_masks.argtypes = [ctypes.c_void_p, _I, _I]
        L.or_set_masks.argtypes = [ctypes.c_void_p, _I, _I]
        L.or_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.or_get_scalars.argtypes = [ctypes.c_void_p, _D]
        for n in ["or_begin_step", "or_explicit_terms_qt", "or_explicit_terms_qr",
                  "or_explicit_terms_qz", "or_implicit_and_update_qt",
                  "or_implicit_and_update_qr", "or_implicit_and_update_qz",
                  "or_calc_local_divergence", "or_solve_pressure_correction",
                  "or_correct_velocity", "or_correct_pressure", "or_time_marcher",
                  "or_project"]:
            getattr(L, n).argtypes = [ctypes.c_void_p]
        L.or_set_substep.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_update_halo.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_calc_max_cfl.restype = ctypes.c_double
        L.or_calc_max_cfl.argtypes = [ctypes.c_void_p]
        L.or_check_divergence.restype = ctypes.c_double
        L.or_check_divergence.argtypes = [ctypes.c_void_p]
        L.or_calc_torque.argtypes = [ctypes.c_void_p, _D, _D]
        L.or_calc_velocity_stats.argtypes = [ctypes.c_void_p, _D, _D]
        L.or_calc_stats.argtypes = [ctypes.c_void_p]
        L.or_calc_dissipation.restype = ctypes.c_double
        L.or_calc_dissipation.argtypes = [ctypes.c_void_p]
        L.or_get_stats.restype = ctypes.c_int
        L.or_get_stats.argtypes = [ctypes.c_void_p, _D, ctypes.POINTER(ctypes.c_int)]
        L.or_last_nute.restype = ctypes.c_double
        L.or_last_nute.argtypes = [ctypes.c_void_p]
        L.or_regrid.restype = ctypes.c_int
        L.or_regrid.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_double, _D]
        L.or_dump_rcut.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D]
        L.or_dump_thcut.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.or_dump_streamwise_avg.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.or_stream_triad.restype = ctypes.c_double
        L.or_stream_triad.argtypes = [ctypes.c_longlong, ctypes.c_int]
        L.or_step.restype = ctypes.c_int
        L.or_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _D]
        for n in ["or_dgttrf", "or_zgttrf"]:
            getattr(L, n).restype = ctypes.c_int
            getattr(L, n).argtypes = [ctypes.c_int, _D, _D, _D, _D, _I]
        for n in ["or_dgttrs", "or_zgttrs"]:
            getattr(L, n).argtypes = [ctypes.c_int, _D, _D, _D, _D, _I, _D]
        L.or_fft_c2c.argtypes = [ctypes.c_int, _D, ctypes.c_int]
        L.or_fft_r2c.argtypes = [ctypes.c_int, _D, _D]
        L.or_fft_c2r.argtypes = [ctypes.c_int, _D, _D]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Oracle:
    """One reference-state instance (module globals of param/local_arrays)."""

    def __init__(self, params):
        self.params = dict(params)
        self._p = make_params(self.params)
        L = lib()
        self.h = L.or_create(ctypes.byref(self._p))
        if not self.h:
            raise ValueError(L.or_last_error().decode())
        self.nthm, self.nzm, self.nrm = params["nthm"], params["nzm"], params["nrm"]
        self.nr = self.nrm + 1
        self.shape = (self.nr, self.nzm, self.nthm)

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- access ---
    def vec(self, name):
        n = 64 if name == "fs" else max(self.nr, self.nthm, self.nzm)
        out = np.zeros(n)
        m = lib().or_get_vec(self.h, VECS.index(name), _dp(out), n)
        return out[:m].copy()

    def get(self, name):
        out = np.zeros(self.shape, order="F")
        lib().or_get_field(self.h, FIELDS[name], _dp(out))
        return out

    def set(self, name, arr):
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(self.shape, order="F"))
        lib().or_set_field(self.h, FIELDS[name], _dp(a))

    def masks(self):
        i = np.zeros((self.nzm, self.nthm), dtype=np.int32, order="F")
        o = np.zeros((self.nzm, self.nthm), dtype=np.int32, order="F")
        lib().or_get_masks(self.h, _ip(i), _ip(o))
        return i, o

    def set_masks(self, icyl, ocyl):
        i = np.asfortranarray(icyl, dtype=np.int32)
        o = np.asfortranarray(ocyl, dtype=np.int32)
        lib().or_set_masks(self.h, _ip(i), _ip(o))

    def set_time(self, time, dt):
        lib().or_set_time(self.h, time, dt)

    def scalars(self):
        o = np.zeros(8)
        lib().or_get_scalars(self.h, _dp(o))
        return dict(zip(["time", "dt", "beta", "al", "ga", "ro", "vinner_t", "dvinner_t"], o))

    # --- reference routines ---
    def begin_step(self):
        lib().or_begin_step(self.h)

    def set_substep(self, ns):
        lib().or_set_substep(self.h, ns)

    def call(self, routine):
        getattr(lib(), "or_" + routine)(self.h)

    def update_halo(self, name):
        lib().or_update_halo(self.h, FIELDS[name])

    def max_cfl(self):
        return lib().or_calc_max_cfl(self.h)

    def max_div(self):
        return lib().or_check_divergence(self.h)

    def torque(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().or_calc_torque(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def velocity_stats(self):
        v, r = np.zeros(3), np.zeros(3)
        lib().or_calc_velocity_stats(self.h, _dp(v), _dp(r))
        return v, r

    def calc_stats(self):
        lib().or_calc_stats(self.h)

    def calc_dissipation(self):
        return lib().or_calc_dissipation(self.h)

    def last_nute(self):
        return lib().or_last_nute(self.h)

    def stats(self):
        """dict name -> (nrm, nzm) accumulator, and the sample count"""
        out = np.zeros((len(STATS), self.nrm, self.nzm), order="C")
        buf = np.zeros(len(STATS) * self.nrm * self.nzm)
        n = ctypes.c_int()
        lib().or_get_stats(self.h, _dp(buf), ctypes.byref(n))
        for q, name in enumerate(STATS):
            out[q] = buf[q * self.nrm * self.nzm:(q + 1) * self.nrm * self.nzm].reshape(
                (self.nrm, self.nzm), order="F")
        return {name: out[q] for q, name in enumerate(STATS)}, n.value

    def regrid(self, name, old, istro, stro):
        """interp.F90 of one field from an old grid (nro, nzo-1, ntho-1) onto
        this grid; returns (nr, nzm, nthm)"""
        old = np.asfortranarray(old, dtype=np.float64)
        nro, nzom, nthom = old.shape
        out = np.zeros(self.shape, order="F")
        rc = lib().or_regrid(self.h, {"qt": 1, "qr": 2, "qz": 3}[name], _dp(old), nthom + 1,
                             nzom + 1, nro, int(istro), float(stro), _dp(out))
        if rc:
            raise ValueError(lib().or_last_error().decode())
        return out

    def dump_rcut(self, kslab):
        """vt, vr, vz at r-plane kslab, each (nzm, nthm)"""
        out = [np.zeros((self.nzm, self.nthm), order="F") for _ in range(3)]
        lib().or_dump_rcut(self.h, int(kslab), *[_dp(a) for a in out])
        return out

    def dump_thcut(self):
        out = [np.zeros((self.nrm, self.nzm), order="F") for _ in range(3)]
        lib().or_dump_thcut(self.h, *[_dp(a) for a in out])
        return out

    def dump_streamwise_avg(self):
        out = [np.zeros((self.nrm, self.nzm), order="F") for _ in range(3)]
        lib().or_dump_streamwise_avg(self.h, *[_dp(a) for a in out])
        return out

    def step(self, ntime, walltime=0.0):
        o = np.zeros(9)
        code = lib().or_step(self.h, ntime, walltime, _dp(o))
        return code, dict(zip(["cfl", "dt", "time", "dmax", "torque_in", "torque_out",
                               "vmax1", "vmax2", "vmax3"], o))


def stream_triad_gbs(n=1 << 26, reps=5):
    """host STREAM triad bandwidth (GB/s) on the OpenMP threads"""
    return lib().or_stream_triad(int(n), int(reps))


# ---- LAPACK / FFT restatements, numpy in and out ----

def dgttrf(dl, d, du):
    n = len(d)
    dl, d, du = (np.array(x, dtype=np.float64) for x in (dl, d, du))
    du2 = np.zeros(max(n - 2, 1))
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().or_dgttrf(n, _dp(dl), _dp(d), _dp(du), _dp(du2), _ip(ipiv))
    return dl, d, du, du2[: max(n - 2, 0)], ipiv, info


def dgttrs(dl, d, du, du2, ipiv, b):
    n = len(d)
    b = np.array(b, dtype=np.float64)
    du2p = np.zeros(max(n - 2, 1))
    du2p[: len(du2)] = du2
    lib().or_dgttrs(n, _dp(np.ascontiguousarray(dl, dtype=np.float64)),
                    _dp(np.ascontiguousarray(d, dtype=np.float64)),
                    _dp(np.ascontiguousarray(du, dtype=np.float64)), _dp(du2p),
                    _ip(np.ascontiguousarray(ipiv, dtype=np.int32)), _dp(b))
    return b


def _c2i(z):
    z = np.asarray(z, dtype=np.complex128)
    out = np.empty(2 * len(z))
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def _i2c(a):
    return a[0::2] + 1j * a[1::2]


def zgttrf(dl, d, du):
    n = len(d)
    dli, di, dui = _c2i(dl), _c2i(d), _c2i(du)
    du2 = np.zeros(2 * max(n - 2, 1))
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().or_zgttrf(n, _dp(dli), _dp(di), _dp(dui), _dp(du2), _ip(ipiv))
    return _i2c(dli), _i2c(di), _i2c(dui), _i2c(du2)[: max(n - 2, 0)], ipiv, info


def zgttrs(dl, d, du, du2, ipiv, b):
    n = len(d)
    du2p = np.zeros(2 * max(n - 2, 1))
    du2p[: 2 * len(du2)] = _c2i(du2)
    bi = _c2i(b)
    lib().or_zgttrs(n, _dp(_c2i(dl)), _dp(_c2i(d)), _dp(_c2i(du)), _dp(du2p),
                    _ip(np.ascontiguousarray(ipiv, dtype=np.int32)), _dp(bi))
    return _i2c(bi)


def fft_c2c(x, sign):
    a = _c2i(x)
    lib().or_fft_c2c(len(x), _dp(a), sign)
    return _i2c(a)


def fft_r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    out = np.zeros(2 * (n // 2 + 1))
    lib().or_fft_r2c(n, _dp(x), _dp(out))
    return _i2c(out)


def fft_c2r(X, n):
    a = _c2i(X)
    out = np.zeros(n)
    lib().or_fft_c2r(n, _dp(a), _dp(out))
    return out
