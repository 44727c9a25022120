"""Python view of libtcamd.so, the MI355X Taylor-Couette RK3 time-step library.

The C-ABI (include/tc_abi.h) is the product boundary; this module is a thin
ctypes binding used by the tests and the benchmark.  `Solver` exposes the
reference's routine names (ExplicitTermsQT, ImplicitAndUpdateQR,
SolvePressureCorrection, ...) on a device-resident context.

Fields are exchanged as numpy arrays of shape (nr, nzm, nthm) in Fortran
order, the reference's (1:nr,1:nzm,1:nthm) without halos.  Importing works
without a GPU; creating a Solver needs one and raises if none is visible or
if the library has not been built.
"""
import ctypes
import os
import struct

import numpy as np

from . import configs  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtcamd.so")

FIELDS = {"qt": 0, "qr": 1, "qz": 2, "pr": 3, "dph": 4, "rhs": 5, "dq": 6,
          "qcap": 7, "ru1": 8, "ru2": 9, "ru3": 10}
VECS = ["rc", "rm", "g2rc", "g2rm", "usrc", "usrm", "udx2rm", "udx2m", "udx2c",
        "rudx2c", "udx1vis1", "udx1vis2", "udx1qm", "ap1j", "am1j", "ac1j",
        "ap2j", "am2j", "ac2j", "ap2je", "am2je", "ac2je", "ap3j", "am3j",
        "ac3j", "apph", "amph", "acph", "ak1", "ak3", "fs", "pvol"]
LAYOUT_FORTRAN, LAYOUT_DEVICE = 0, 1

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

# every entry point declared in include/tc_abi.h
ABI_SYMBOLS = [
    "tc_create", "tc_destroy", "tc_last_error", "tc_grid_vector", "tc_upload_field",
    "tc_download_field", "tc_set_masks", "tc_get_masks", "tc_set_time", "tc_get_scalars",
    "tc_begin_step", "tc_set_substep", "tc_explicit_terms_qt", "tc_explicit_terms_qr",
    "tc_explicit_terms_qz", "tc_implicit_and_update_qt", "tc_implicit_and_update_qr",
    "tc_implicit_and_update_qz", "tc_calc_local_divergence", "tc_solve_pressure_correction",
    "tc_correct_velocity", "tc_correct_pressure", "tc_time_marcher", "tc_set_fused",
    "tc_calc_max_cfl", "tc_check_divergence", "tc_calc_torque", "tc_calc_velocity_stats",
    "tc_step", "tc_project", "tc_init_perturbed_couette", "tc_synchronize",
    "tc_kernel_timing", "tc_kernel_timing_read", "tc_copy_field", "tc_device_bytes",
    "tc_create_tuned", "tc_set_tuning", "tc_get_tuning", "tc_autotune",
    "tc_batch_create", "tc_batch_destroy", "tc_batch_add", "tc_batch_remove", "tc_batch_ctx",
    "tc_batch_advance", "tc_batch_status", "tc_batch_groups",
    "tc_batch_write_restart", "tc_batch_read_restart",
    "tc_write_restart", "tc_read_restart",
    "tc_calc_stats", "tc_calc_dissipation", "tc_get_stats", "tc_set_stats",
    "tc_last_dissipation", "tc_dump_rcut", "tc_dump_thcut", "tc_dump_streamwise_avg",
    "tc_regrid_field", "tc_stream_triad", "tc_kernel_timing_only", "tc_run_steps",
    "tc_run_steps_logged", "tc_check_divergence_noalloc",
]

# stat_arrays in WriteStats order (StatRoutines.F90:97-152); TC_ST_* ids
STATS = ["vt_mean", "vr_mean", "vz_mean", "wt_mean", "wr_mean", "wz_mean",
         "vt_rms", "vr_rms", "vz_rms", "wt_rms", "wr_rms", "wz_rms",
         "vtvr_mean", "vtvz_mean", "vrvz_mean", "pr_mean", "pr_rms", "disste"]

TUNING_FIELDS = ["helm_lpt", "fft_t_lds", "fft_z_lds", "fft_z_nt", "fft_pad", "poisson_pq",
                 "defer"]


class Params(ctypes.Structure):
    _fields_ = PARAM_FIELDS


def make_params(d):
    p = Params()
    for name, _ in PARAM_FIELDS:
        setattr(p, name, d.get(name, 0))
    return p


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in TUNING_FIELDS]


def make_tuning(d):
    t = Tuning()
    for n in TUNING_FIELDS:
        setattr(t, n, int((d or {}).get(n, 0)))
    return t


def tuning_dict(t):
    return {n: getattr(t, n) for n in TUNING_FIELDS}


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p
_LIB = None


class TCError(RuntimeError):
    pass


def lib():
    """Loads libtcamd.so; raises if it has not been built (no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise TCError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                      "or python taylor-couette-afid_amd/build.py")
    L = ctypes.CDLL(LIBPATH)
    L.tc_last_error.restype = ctypes.c_char_p
    L.tc_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(_V)]
    L.tc_destroy.argtypes = [_V]
    L.tc_grid_vector.argtypes = [ctypes.POINTER(Params), ctypes.c_int, _D, ctypes.c_int]
    L.tc_upload_field.argtypes = [_V, ctypes.c_int, _D, ctypes.c_int]
    L.tc_download_field.argtypes = [_V, ctypes.c_int, _D, ctypes.c_int]
    L.tc_set_masks.argtypes = [_V, _I, _I]
    L.tc_get_masks.argtypes = [_V, _I, _I]
    L.tc_set_time.argtypes = [_V, ctypes.c_double, ctypes.c_double]
    L.tc_get_scalars.argtypes = [_V, _D]
    L.tc_set_substep.argtypes = [_V, ctypes.c_int]
    L.tc_set_fused.argtypes = [_V, ctypes.c_int]
    for n in ["tc_begin_step", "tc_explicit_terms_qt", "tc_explicit_terms_qr",
              "tc_explicit_terms_qz", "tc_implicit_and_update_qt", "tc_implicit_and_update_qr",
              "tc_implicit_and_update_qz", "tc_calc_local_divergence",
              "tc_solve_pressure_correction", "tc_correct_velocity", "tc_correct_pressure",
              "tc_time_marcher", "tc_project", "tc_synchronize"]:
        getattr(L, n).argtypes = [_V]
    L.tc_calc_max_cfl.argtypes = [_V, _D]
    L.tc_check_divergence.argtypes = [_V, _D]
    L.tc_check_divergence_noalloc.argtypes = [_V, ctypes.c_void_p]
    L.tc_calc_torque.argtypes = [_V, _D, _D]
    L.tc_calc_velocity_stats.argtypes = [_V, _D, _D]
    L.tc_step.argtypes = [_V, ctypes.c_int, ctypes.c_double, _D]
    L.tc_run_steps.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I, _D]
    L.tc_run_steps_logged.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I, _D,
                                      _D, ctypes.c_int, _I]
    L.tc_init_perturbed_couette.argtypes = [ctypes.POINTER(Params), ctypes.c_ulonglong,
                                            ctypes.c_double, _D, _D, _D]
    L.tc_kernel_timing.argtypes = [_V, ctypes.c_int]
    L.tc_kernel_timing_only.argtypes = [_V, ctypes.c_char_p]
    L.tc_kernel_timing_read.argtypes = [_V, ctypes.c_char_p, _D,
                                        ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    L.tc_device_bytes.argtypes = [_V, ctypes.POINTER(ctypes.c_ulonglong)]
    L.tc_copy_field.argtypes = [_V, ctypes.c_int, ctypes.c_int]
    L.tc_stream_triad.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    _T = ctypes.POINTER(Tuning)
    L.tc_create_tuned.argtypes = [ctypes.POINTER(Params), _T, ctypes.POINTER(_V)]
    L.tc_set_tuning.argtypes = [_V, _T]
    L.tc_get_tuning.argtypes = [_V, _T]
    L.tc_autotune.argtypes = [ctypes.POINTER(Params), ctypes.c_int, _T]
    L.tc_batch_create.argtypes = [ctypes.POINTER(_V)]
    L.tc_batch_destroy.argtypes = [_V]
    L.tc_batch_add.argtypes = [_V, ctypes.POINTER(Params), _T, _I]
    L.tc_batch_remove.argtypes = [_V, ctypes.c_int]
    L.tc_batch_ctx.argtypes = [_V, ctypes.c_int, ctypes.POINTER(_V)]
    L.tc_batch_advance.argtypes = [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.tc_batch_status.argtypes = [_V, ctypes.c_int, _I, _I, _D]
    L.tc_batch_groups.argtypes = [_V, _I, _I]
    L.tc_batch_write_restart.argtypes = [_V, ctypes.c_int, ctypes.c_char_p]
    L.tc_batch_read_restart.argtypes = [_V, ctypes.c_int, ctypes.c_char_p, _I]
    L.tc_write_restart.argtypes = [_V, ctypes.c_char_p, ctypes.c_int]
    L.tc_read_restart.argtypes = [_V, ctypes.c_char_p, _I]
    L.tc_calc_stats.argtypes = [_V]
    L.tc_calc_dissipation.argtypes = [_V, _D]
    L.tc_get_stats.argtypes = [_V, _D, _I]
    L.tc_set_stats.argtypes = [_V, _D, ctypes.c_int]
    L.tc_last_dissipation.argtypes = [_V, _D]
    L.tc_dump_rcut.argtypes = [_V, ctypes.c_int, _D, _D, _D]
    L.tc_regrid_field.argtypes = [_V, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double]
    L.tc_dump_thcut.argtypes = [_V, _D, _D, _D]
    L.tc_dump_streamwise_avg.argtypes = [_V, _D, _D, _D]
    _LIB = L
    return L


def _check(rc):
    if rc < 0:
        raise TCError(lib().tc_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(_D)


def grid_vector(params, name):
    """CreateGrid / InitPressureSolver vector computed by the library (host only)."""
    p = make_params(params)
    n = max(params["nrm"] + 1, params["nthm"], params["nzm"], 16)
    out = np.zeros(n)
    m = _check(lib().tc_grid_vector(ctypes.byref(p), VECS.index(name), _dp(out), n))
    return out[:m].copy()


def init_perturbed_couette(params, seed, amp):
    """Seeded starting velocity (qt, qr, qz), each (nr, nzm, nthm) Fortran order."""
    p = make_params(params)
    shape = (params["nrm"] + 1, params["nzm"], params["nthm"])
    qt, qr, qz = (np.zeros(shape, order="F") for _ in range(3))
    _check(lib().tc_init_perturbed_couette(ctypes.byref(p), seed, amp, _dp(qt), _dp(qr), _dp(qz)))
    return qt, qr, qz


class Solver:
    """One device-resident Taylor-Couette instance with the AFiD routine names."""

    def __init__(self, params, tuning=None, _handle=None):
        self.params = dict(params)
        self._p = make_params(self.params)
        self._owned = _handle is None
        if _handle is None:
            h = _V()
            if tuning is None:
                _check(lib().tc_create(ctypes.byref(self._p), ctypes.byref(h)))
            else:
                t = make_tuning(tuning)
                _check(lib().tc_create_tuned(ctypes.byref(self._p), ctypes.byref(t), ctypes.byref(h)))
            self.h = h
        else:
            self.h = _handle
        self.nthm, self.nzm, self.nrm = params["nthm"], params["nzm"], params["nrm"]
        self.nr = self.nrm + 1
        self.shape = (self.nr, self.nzm, self.nthm)

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                lib().tc_destroy(self.h)
            self.h = None

    def tuning(self):
        t = Tuning()
        _check(lib().tc_get_tuning(self.h, ctypes.byref(t)))
        return tuning_dict(t)

    def set_tuning(self, tuning):
        t = make_tuning(tuning)
        _check(lib().tc_set_tuning(self.h, ctypes.byref(t)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state ---
    def set(self, name, arr):
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(self.shape, order="F"))
        _check(lib().tc_upload_field(self.h, FIELDS[name], _dp(a), LAYOUT_FORTRAN))

    def get(self, name):
        out = np.zeros(self.shape, order="F")
        _check(lib().tc_download_field(self.h, FIELDS[name], _dp(out), LAYOUT_FORTRAN))
        return out

    def set_masks(self, icyl, ocyl):
        i = np.asfortranarray(icyl, dtype=np.int32)
        o = np.asfortranarray(ocyl, dtype=np.int32)
        _check(lib().tc_set_masks(self.h, i.ctypes.data_as(_I), o.ctypes.data_as(_I)))

    def masks(self):
        i = np.zeros((self.nzm, self.nthm), dtype=np.int32, order="F")
        o = np.zeros((self.nzm, self.nthm), dtype=np.int32, order="F")
        _check(lib().tc_get_masks(self.h, i.ctypes.data_as(_I), o.ctypes.data_as(_I)))
        return i, o

    def set_time(self, time, dt):
        _check(lib().tc_set_time(self.h, time, dt))

    def scalars(self):
        o = np.zeros(8)
        _check(lib().tc_get_scalars(self.h, _dp(o)))
        return dict(zip(["time", "dt", "beta", "al", "ga", "ro", "vinner_t", "dvinner_t"], o))

    def set_fused(self, on):
        _check(lib().tc_set_fused(self.h, int(bool(on))))

    def synchronize(self):
        _check(lib().tc_synchronize(self.h))

    # --- reference routines (TimeMarcher.F90:20-54 call order) ---
    def begin_step(self):
        _check(lib().tc_begin_step(self.h))

    def set_substep(self, ns):
        _check(lib().tc_set_substep(self.h, ns))

    def ExplicitTermsQT(self):
        _check(lib().tc_explicit_terms_qt(self.h))

    def ExplicitTermsQR(self):
        _check(lib().tc_explicit_terms_qr(self.h))

    def ExplicitTermsQZ(self):
        _check(lib().tc_explicit_terms_qz(self.h))

    def ImplicitAndUpdateQT(self):
        _check(lib().tc_implicit_and_update_qt(self.h))

    def ImplicitAndUpdateQR(self):
        _check(lib().tc_implicit_and_update_qr(self.h))

    def ImplicitAndUpdateQZ(self):
        _check(lib().tc_implicit_and_update_qz(self.h))

    def CalcLocalDivergence(self):
        _check(lib().tc_calc_local_divergence(self.h))

    def SolvePressureCorrection(self):
        _check(lib().tc_solve_pressure_correction(self.h))

    def CorrectVelocity(self):
        _check(lib().tc_correct_velocity(self.h))

    def CorrectPressure(self):
        _check(lib().tc_correct_pressure(self.h))

    def TimeMarcher(self):
        _check(lib().tc_time_marcher(self.h))

    def CalcMaxCFL(self):
        v = ctypes.c_double()
        _check(lib().tc_calc_max_cfl(self.h, ctypes.byref(v)))
        return v.value

    def CheckDivergence(self):
        v = ctypes.c_double()
        _check(lib().tc_check_divergence(self.h, ctypes.byref(v)))
        return v.value

    def CalcTorque(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().tc_calc_torque(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def CalcVelocityStats(self):
        v, r = np.zeros(3), np.zeros(3)
        _check(lib().tc_calc_velocity_stats(self.h, _dp(v), _dp(r)))
        return v, r

    def project(self):
        _check(lib().tc_project(self.h))

    def write_restart(self, path, ntime):
        """Whole-state checkpoint (fields, pressure, RK histories, time);
        restart_info reads its header."""
        _check(lib().tc_write_restart(self.h, os.fsencode(path), int(ntime)))

    def read_restart(self, path):
        """Loads a checkpoint; returns the ntime it was written at."""
        n = ctypes.c_int()
        _check(lib().tc_read_restart(self.h, os.fsencode(path), ctypes.byref(n)))
        return n.value

    # --- statistics (StatRoutines.F90, CalcVorticity.F90, CalcDissipation.F90) ---
    def CalcStats(self):
        _check(lib().tc_calc_stats(self.h))

    def CalcDissipation(self):
        """Adds to disste; returns nute."""
        v = np.zeros(1)
        _check(lib().tc_calc_dissipation(self.h, _dp(v)))
        return float(v[0])

    def last_dissipation(self):
        v = np.zeros(1)
        _check(lib().tc_last_dissipation(self.h, _dp(v)))
        return float(v[0])

    def stats(self):
        """dict name -> (nrm, nzm) accumulator, and the sample count"""
        m = self.nrm * self.nzm
        buf = np.zeros(len(STATS) * m)
        n = ctypes.c_int()
        _check(lib().tc_get_stats(self.h, _dp(buf), ctypes.byref(n)))
        return {name: buf[q * m:(q + 1) * m].reshape((self.nrm, self.nzm), order="F")
                for q, name in enumerate(STATS)}, n.value

    def set_stats(self, stats, nsamples):
        m = self.nrm * self.nzm
        buf = np.zeros(len(STATS) * m)
        for q, name in enumerate(STATS):
            buf[q * m:(q + 1) * m] = np.asarray(stats[name], dtype=np.float64).reshape(-1, order="F")
        _check(lib().tc_set_stats(self.h, _dp(buf), int(nsamples)))

    def regrid(self, name, old, istro, stro):
        """ReadFlowField's interpolation (interp.F90) of one field from an old
        grid, old shaped (nro, nzo-1, ntho-1), into this context."""
        old = np.asfortranarray(old, dtype=np.float64)
        nro, nzom, nthom = old.shape
        _check(lib().tc_regrid_field(self.h, FIELDS[name], _dp(old), nthom + 1, nzom + 1, nro,
                                     int(istro), float(stro)))

    # --- slab dumps (DumpRCuts.F90, DumpThCuts.F90, DumpStreamwiseAvg.F90) ---
    def DumpRCut(self, kslab):
        """vt, vr, vz at r-plane kslab (1..nrm), each (nzm, nthm)"""
        out = [np.zeros((self.nzm, self.nthm), order="F") for _ in range(3)]
        _check(lib().tc_dump_rcut(self.h, int(kslab), *[_dp(a) for a in out]))
        return out

    def DumpThCut(self):
        out = [np.zeros((self.nrm, self.nzm), order="F") for _ in range(3)]
        _check(lib().tc_dump_thcut(self.h, *[_dp(a) for a in out]))
        return out

    def DumpStreamwiseAvg(self):
        out = [np.zeros((self.nrm, self.nzm), order="F") for _ in range(3)]
        _check(lib().tc_dump_streamwise_avg(self.h, *[_dp(a) for a in out]))
        return out

    def step(self, ntime, walltime=0.0):
        """One main.F90 loop iteration; returns (errorcode, diagnostics)."""
        o = np.zeros(9)
        code = _check(lib().tc_step(self.h, ntime, walltime, _dp(o)))
        return code, dict(zip(["cfl", "dt", "time", "dmax", "torque_in", "torque_out",
                               "vmax1", "vmax2", "vmax3"], o))

    def run(self, ntime0, nsteps, walltime=0.0):
        """tc_step from ntime0 for at most nsteps iterations, stopping at the
        first error code; returns (errorcode, last ntime, diagnostics)."""
        o = np.zeros(9)
        n = ctypes.c_int()
        code = _check(lib().tc_run_steps(self.h, int(ntime0), int(nsteps), float(walltime),
                                         ctypes.byref(n), _dp(o)))
        return code, n.value, dict(zip(["cfl", "dt", "time", "dmax", "torque_in", "torque_out",
                                        "vmax1", "vmax2", "vmax3"], o))

    def run_logged(self, ntime0, nsteps, walltime=0.0, log_cap=None):
        """run() that also returns one (ntime, diagnostics) row per logging
        iteration, as the reference's logs record them."""
        cap = nsteps if log_cap is None else int(log_cap)
        o = np.zeros(9)
        rows = np.zeros((max(cap, 1), 10))
        n, nl = ctypes.c_int(), ctypes.c_int()
        code = _check(lib().tc_run_steps_logged(self.h, int(ntime0), int(nsteps), float(walltime),
                                                ctypes.byref(n), _dp(o), _dp(rows), cap,
                                                ctypes.byref(nl)))
        keys = ["cfl", "dt", "time", "dmax", "torque_in", "torque_out", "vmax1", "vmax2", "vmax3"]
        log = [(int(r[0]), dict(zip(keys, r[1:]))) for r in rows[:min(nl.value, cap)]]
        return code, n.value, dict(zip(keys, o)), log

    # --- instrumentation ---
    def kernel_timing(self, on, only=None):
        """per-kernel HIP-event timing; only: time just that kernel name"""
        if only:
            _check(lib().tc_kernel_timing_only(self.h, only.encode()))
        else:
            _check(lib().tc_kernel_timing(self.h, int(bool(on))))

    def kernel_time(self, name, reset=False):
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        _check(lib().tc_kernel_timing_read(self.h, name.encode(), ctypes.byref(ms),
                                           ctypes.byref(cnt), int(bool(reset))))
        return ms.value, cnt.value

    def copy_field(self, dst, src):
        _check(lib().tc_copy_field(self.h, FIELDS[dst], FIELDS[src]))

    def stream_triad(self, a, b, c, s=3.0):
        _check(lib().tc_stream_triad(self.h, FIELDS[a], FIELDS[b], FIELDS[c], float(s)))

    def device_bytes(self):
        v = ctypes.c_ulonglong()
        _check(lib().tc_device_bytes(self.h, ctypes.byref(v)))
        return v.value


_RESTART_HEADER = struct.Struct("<8s8i8d")


def restart_info(path):
    """The header of a checkpoint written by tc_write_restart (8-byte magic,
    int32 nthm nzm nrm nsst ntime cfl_valid istr pad, float64 time dt cfl
    rint rext alx3 lamb str), without reading its fields."""
    with open(path, "rb") as f:
        raw = f.read(_RESTART_HEADER.size)
    if len(raw) != _RESTART_HEADER.size or raw[:8] != b"TCAFIDR1":
        raise TCError(f"{path} is not a restart file of this library")
    v = _RESTART_HEADER.unpack(raw)
    return dict(nthm=v[1], nzm=v[2], nrm=v[3], nsst=v[4], ntime=v[5], time=v[9], dt=v[10])


def autotune(params, steps=3):
    """Fastest schedule for this grid shape (timed on the device, cached per
    shape for the life of the process)."""
    p = make_params(params)
    t = Tuning()
    _check(lib().tc_autotune(ctypes.byref(p), int(steps), ctypes.byref(t)))
    return tuning_dict(t)


class Batch:
    """Many independent instances advanced together with grouped launches
    (one launch per kernel per shape group) and an optional hipGraph per
    step.  Instances advance at a fixed dt (IDTV = 0, dt = DTMAX)."""

    STATUS = ["cfl", "dt", "time", "dmax", "torque_in", "torque_out", "vmax1", "vmax2", "vmax3"]

    def __init__(self):
        h = _V()
        _check(lib().tc_batch_create(ctypes.byref(h)))
        self.h = h
        self.params = {}

    def close(self):
        if getattr(self, "h", None):
            lib().tc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, params, tuning=None):
        p = make_params(params)
        i = ctypes.c_int()
        if tuning is None:
            _check(lib().tc_batch_add(self.h, ctypes.byref(p), None, ctypes.byref(i)))
        else:
            t = make_tuning(tuning)
            _check(lib().tc_batch_add(self.h, ctypes.byref(p), ctypes.byref(t), ctypes.byref(i)))
        self.params[i.value] = dict(params)
        return i.value

    def remove(self, iid):
        _check(lib().tc_batch_remove(self.h, int(iid)))
        self.params.pop(iid, None)

    def solver(self, iid):
        """Solver view of one instance (not owned: the batch frees it)."""
        h = _V()
        _check(lib().tc_batch_ctx(self.h, int(iid), ctypes.byref(h)))
        return Solver(self.params[iid], _handle=h)

    def advance(self, nsteps, graph=True, sync_every=16):
        return _check(lib().tc_batch_advance(self.h, int(nsteps), int(bool(graph)), int(sync_every)))

    def status(self, iid):
        steps, code = ctypes.c_int(), ctypes.c_int()
        o = np.zeros(9)
        _check(lib().tc_batch_status(self.h, int(iid), ctypes.byref(steps), ctypes.byref(code), _dp(o)))
        return steps.value, code.value, dict(zip(self.STATUS, o))

    def write_restart(self, iid, path):
        """Whole-state checkpoint of one instance at its current step."""
        _check(lib().tc_batch_write_restart(self.h, int(iid), os.fsencode(path)))

    def read_restart(self, iid, path):
        """Loads a checkpoint into an instance that has not stepped yet;
        returns the step it continues from."""
        n = ctypes.c_int()
        _check(lib().tc_batch_read_restart(self.h, int(iid), os.fsencode(path), ctypes.byref(n)))
        return n.value

    def groups(self):
        n, l = ctypes.c_int(), ctypes.c_int()
        _check(lib().tc_batch_groups(self.h, ctypes.byref(n), ctypes.byref(l)))
        return n.value, l.value
