"""Routine-by-routine parity of the HIP path against the oracle.

Each test loads the same random state into a device context and into the
CPU restatement, calls one reference routine on both through their C
interfaces, and compares every field the routine writes.  Bar: relative L2
<= 1e-12 (FP64 stencils and solves; the two sides differ only by FMA
contraction and reciprocal-versus-division rounding).
"""
import numpy as np
import pytest

from helpers import load_both, random_state, rel_l2, small_params

pytestmark = pytest.mark.gpu
TOL = 1e-12

SHAPES = [(32, 24, 33), (64, 64, 65), (48, 40, 17)]


def _setup(tc, oracle, shape, seed, masks=False, ns=2, time=0.37, dt=2e-3, ampl=0.05):
    nthm, nzm, nrm = shape
    p = small_params(tc, nthm, nzm, nrm, sizepert=ampl, omegapert=0.7, romeg=0.3,
                     virotext=-0.2)
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    st = random_state(p, seed)
    load_both(dev, orc, st)
    for s in (dev, orc):
        s.set_time(time, dt)
        s.begin_step()
        s.set_substep(ns)
    if masks:
        # after begin_step: TimeMarcher's SlipBCs call would reset them
        rng = np.random.default_rng(seed + 7)
        ic = rng.integers(0, 2, (nzm, nthm)).astype(np.int32)
        oc = rng.integers(0, 2, (nzm, nthm)).astype(np.int32)
        dev.set_masks(ic, oc)
        orc.set_masks(ic, oc)
    return p, dev, orc


def _cmp(dev, orc, name, planes, tol=TOL):
    a = dev.get(name)[planes]
    b = orc.get(name)[planes]
    e = rel_l2(a, b)
    assert e <= tol, f"{name}: rel L2 {e:.3e}"
    return e


@pytest.mark.parametrize("shape", SHAPES)
def test_explicit_terms(tc, oracle, shape):
    p, dev, orc = _setup(tc, oracle, shape, 11)
    nrm = shape[2]
    dev.ExplicitTermsQT(); orc.call("explicit_terms_qt")
    dev.ExplicitTermsQR(); orc.call("explicit_terms_qr")
    dev.ExplicitTermsQZ(); orc.call("explicit_terms_qz")
    _cmp(dev, orc, "dq", slice(0, nrm))
    _cmp(dev, orc, "dph", slice(1, nrm))
    _cmp(dev, orc, "qcap", slice(0, nrm))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("masks", [False, True])
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_implicit_and_update(tc, oracle, shape, masks, ns):
    p, dev, orc = _setup(tc, oracle, shape, 21 + ns, masks=masks, ns=ns)
    nrm = shape[2]
    dev.ImplicitAndUpdateQT(); orc.call("implicit_and_update_qt")
    dev.ImplicitAndUpdateQR(); orc.call("implicit_and_update_qr")
    dev.ImplicitAndUpdateQZ(); orc.call("implicit_and_update_qz")
    _cmp(dev, orc, "qt", slice(0, nrm))
    _cmp(dev, orc, "ru1", slice(0, nrm))
    _cmp(dev, orc, "qr", slice(0, nrm + 1))
    _cmp(dev, orc, "ru2", slice(1, nrm))
    _cmp(dev, orc, "qz", slice(0, nrm))
    _cmp(dev, orc, "ru3", slice(0, nrm))


@pytest.mark.parametrize("shape", SHAPES)
def test_divergence_and_corrections(tc, oracle, shape):
    p, dev, orc = _setup(tc, oracle, shape, 31)
    nrm = shape[2]
    dev.CalcLocalDivergence(); orc.call("calc_local_divergence")
    _cmp(dev, orc, "dph", slice(0, nrm))
    # corrections from a common random dph
    p, dev, orc = _setup(tc, oracle, shape, 32)
    dev.CorrectVelocity(); orc.call("correct_velocity")
    dev.CorrectPressure(); orc.call("correct_pressure")
    _cmp(dev, orc, "qt", slice(0, nrm))
    _cmp(dev, orc, "qr", slice(0, nrm + 1))
    _cmp(dev, orc, "qz", slice(0, nrm))
    _cmp(dev, orc, "pr", slice(0, nrm))


@pytest.mark.parametrize("shape", SHAPES + [(96, 48, 25), (30, 18, 21)])
def test_solve_pressure_correction(tc, oracle, shape):
    p, dev, orc = _setup(tc, oracle, shape, 41)
    nrm = shape[2]
    dev.SolvePressureCorrection(); orc.call("solve_pressure_correction")
    _cmp(dev, orc, "dph", slice(0, nrm), tol=1e-11)


# 26 rows per thread (the default beyond nrm = 512, where 13-row partitions
# make the reduced system long) and 13, on short and long radial lines
@pytest.mark.parametrize("pq", [13, 26])
@pytest.mark.parametrize("shape", [(32, 24, 33), (16, 8, 600), (24, 16, 769)])
def test_solve_pressure_correction_rows_per_thread(tc, oracle, shape, pq):
    p, dev, orc = _setup(tc, oracle, shape, 43)
    nrm = shape[2]
    assert dev.tuning()["poisson_pq"] == (26 if nrm > 512 else 13)
    dev.set_tuning({"poisson_pq": pq})
    dev.SolvePressureCorrection(); orc.call("solve_pressure_correction")
    _cmp(dev, orc, "dph", slice(0, nrm), tol=1e-11)


# z-FFT block shapes: 256 and 512 (default) threads with 32 and 64 KiB of
# LDS (column tiles of 1 to 64 ktheta), on radix-2/3/5 lengths
@pytest.mark.parametrize("nt,lds", [(256, 32768), (512, 65536), (512, 32768), (256, 65536)])
@pytest.mark.parametrize("shape", [(32, 24, 9), (40, 1024, 5), (30, 360, 7)])
def test_solve_pressure_correction_z_blocks(tc, oracle, shape, nt, lds):
    p, dev, orc = _setup(tc, oracle, shape, 47)
    nrm = shape[2]
    dev.set_tuning({"fft_z_nt": nt, "fft_z_lds": lds})
    assert dev.tuning()["fft_z_nt"] == nt
    dev.SolvePressureCorrection(); orc.call("solve_pressure_correction")
    _cmp(dev, orc, "dph", slice(0, nrm), tol=1e-11)


@pytest.mark.parametrize("shape", SHAPES)
def test_reductions(tc, oracle, shape):
    p, dev, orc = _setup(tc, oracle, shape, 51, masks=True)
    assert abs(dev.CalcMaxCFL() / orc.max_cfl() - 1) < 1e-14
    assert abs(dev.CheckDivergence() / orc.max_div() - 1) < 1e-13
    ti, to = dev.CalcTorque()
    ri, ro = orc.torque()
    assert abs(ti / ri - 1) < 1e-12 and abs(to / ro - 1) < 1e-12
    v, r = dev.CalcVelocityStats()
    vo, rr = orc.velocity_stats()
    np.testing.assert_allclose(v, vo, rtol=1e-15)
    np.testing.assert_allclose(r, rr, rtol=1e-12)


@pytest.mark.parametrize("strips", [(2, 3, 1), (0, 2, 2), (1, 1, 3)])
def test_slip_bc_masks(tc, oracle, strips):
    """SlipBCs.F90:31-90 patterns at several times, as the device context
    computes them at TimeMarcher's step start."""
    p = small_params(tc, 48, 40, 17)
    p.update(minstrt=strips[0], minstrz=strips[1], multmin=strips[2])
    dev = tc.Solver(p)
    orc = oracle.Oracle(p)
    for t in (0.0, 0.123, 0.9, 2.5):
        for s in (dev, orc):
            s.set_time(t, 1e-3)
            s.begin_step()
        a, b = dev.masks(), orc.masks()
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
