"""The Fortran host (tc_host: bou.in, main loop, AFiD routine names over the
C-ABI) on the device, routine by routine and fused, against the oracle run
of the same bou.in: torque at every step within 1e-9, final fields within
1e-10 relative L2."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_l2, write_bou

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "taylor-couette-afid_amd", "tc_host")


def _params(tc):
    p = tc.configs.scaled(tc.configs.small_64(), 32, 24, 33)
    p.update(ntst=5, tout=1e-9, sizepert=0.1, omegapert=0.5, idtv=1, dtmin=1e-3, dtmax=5e-3,
             limitcfl=0.5, virotext=-0.2)
    return p


def _run_host(tmp, p, fused):
    write_bou(tmp / "bou.in", p)
    cmd = [HOST, "--ic", "perturbed", "77", "0.01"] + (["--fused"] if fused else [])
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "exit code 555" in r.stdout, r.stdout
    tq = np.loadtxt(tmp / "torquenu_full.out", ndmin=2)
    shape = (p["nrm"] + 1, p["nzm"], p["nthm"])
    fields = {n: np.fromfile(tmp / f"continua_{n}.bin").reshape(shape, order="F")
              for n in ("qt", "qr", "qz", "pr")}
    return tq, fields


def _oracle_run(tc, oracle, p):
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 77, 0.01)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    dev.project()
    o = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        o.set(n, dev.get(n))
    dev.close()
    tq = []
    for n in range(p["ntst"] + 1):
        code, d = o.step(n)
        tq.append((d["time"], d["torque_in"], d["torque_out"]))
        if code:
            break
    assert code == 555
    return np.array(tq), {n: o.get(n) for n in ("qt", "qr", "qz", "pr")}


@pytest.mark.parametrize("fused", [False, True])
def test_fortran_host_matches_oracle(tc, oracle, tmp_path, fused):
    if not os.path.exists(HOST):
        pytest.fail("tc_host is not built")
    p = _params(tc)
    tq, f = _run_host(tmp_path, p, fused)
    tq_o, f_o = _oracle_run(tc, oracle, p)
    assert tq.shape == tq_o.shape
    np.testing.assert_allclose(tq[:, 0], tq_o[:, 0], rtol=1e-12)
    np.testing.assert_allclose(tq[:, 1:], tq_o[:, 1:], rtol=1e-9)
    nrm = p["nrm"]
    for n in ("qt", "qz"):
        assert rel_l2(f[n][:nrm], f_o[n][:nrm]) < 1e-10, n
    assert rel_l2(f["qr"], f_o["qr"]) < 1e-10
    d = f["pr"][:nrm] - f_o["pr"][:nrm]
    d -= d.mean(axis=(1, 2), keepdims=True)
    ref = f_o["pr"][:nrm] - f_o["pr"][:nrm].mean(axis=(1, 2), keepdims=True)
    assert np.linalg.norm(d) / np.linalg.norm(ref) < 1e-10


def test_error_codes_match_oracle(tc, oracle):
    """main.F90 exit codes forced from bou.in values, device against oracle."""
    base = tc.configs.scaled(tc.configs.small_64(), 8, 8, 9)
    base.update(ren=100.0, tout=1e30, ntst=50, tmax=1e30)
    cases = [dict(ntst=2), dict(resid=1e-300), dict(tmax=1e-6), dict(limitvel=1e-9),
             dict(idtv=0, dtmax=1e-3, limitcfl=1e-9),
             dict(idtv=1, dtmin=2e-3, dtmax=1e-2, limitcfl=1e-9)]
    for case in cases:
        p = dict(base, **case)
        codes = []
        for side in ("dev", "orc"):
            s = tc.Solver(p) if side == "dev" else oracle.Oracle(p)
            rng = np.random.default_rng(1)
            q = np.zeros((p["nrm"] + 1, p["nzm"], p["nthm"]), order="F")
            q[1:p["nrm"]] = 1e-2 * rng.uniform(-1, 1, (p["nrm"] - 1, p["nzm"], p["nthm"]))
            s.set("qr", q)
            (s.project() if side == "dev" else s.call("project"))
            seq = []
            for k in range(5):
                c, _ = s.step(k)
                seq.append(c)
                if c:
                    break
            codes.append(seq)
        assert codes[0] == codes[1], (case, codes)


def test_fortran_host_checkpoint_restart(tc, tmp_path):
    """tc_host --checkpoint after 3 steps, then --restart for the last 3: the
    torque of every step and the final fields equal the uninterrupted run."""
    if not os.path.exists(HOST):
        pytest.fail("tc_host is not built")
    p = _params(tc)
    full = tmp_path / "full"
    full.mkdir()
    tq_full, f_full = _run_host(full, p, fused=True)
    part = tmp_path / "part"
    part.mkdir()
    write_bou(part / "bou.in", dict(p, ntst=2))
    r = subprocess.run([HOST, "--ic", "perturbed", "77", "0.01", "--fused", "--checkpoint", "ck.bin"],
                       cwd=part, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "exit code 555" in r.stdout, r.stdout + r.stderr
    tq_a = np.loadtxt(part / "torquenu_full.out", ndmin=2)
    write_bou(part / "bou.in", p)
    r = subprocess.run([HOST, "--restart", "ck.bin", "--fused"], cwd=part, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "exit code 555" in r.stdout, r.stdout + r.stderr
    tq_b = np.loadtxt(part / "torquenu_full.out", ndmin=2)
    assert np.array_equal(np.vstack([tq_a, tq_b]), tq_full)
    shape = (p["nrm"] + 1, p["nzm"], p["nthm"])
    for n in ("qt", "qr", "qz", "pr"):
        got = np.fromfile(part / f"continua_{n}.bin").reshape(shape, order="F")
        assert np.array_equal(got, f_full[n]), n


def test_fortran_host_statistics(tc, oracle, tmp_path):
    """TSTA inside the run: CalcStats and CalcDissipation at every logging
    step after it, dissipnu.out, and stats/*.bin written on the normal exit
    match the oracle's accumulators."""
    if not os.path.exists(HOST):
        pytest.fail("tc_host is not built")
    p = _params(tc)
    p.update(tsta=0.004)
    (tmp_path / "stats").mkdir()
    write_bou(tmp_path / "bou.in", p)
    r = subprocess.run([HOST, "--ic", "perturbed", "77", "0.01", "--fused"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "exit code 555" in r.stdout, r.stdout + r.stderr
    dev = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, 77, 0.01)
    dev.set("qt", qt); dev.set("qr", qr); dev.set("qz", qz)
    dev.project()
    o = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        o.set(n, dev.get(n))
    dev.close()
    nute = []
    for n in range(p["ntst"] + 1):
        code, d = o.step(n)
        if d["time"] > p["tsta"]:
            nute.append((d["time"], o.last_nute()))
    assert code == 555 and len(nute) >= 2
    got = np.loadtxt(tmp_path / "dissipnu.out", ndmin=2)
    assert got.shape == (len(nute), 2)
    np.testing.assert_allclose(got, np.array(nute), rtol=1e-9)
    want, ns = o.stats()
    assert int(open(tmp_path / "stats" / "averaging_time.txt").read().split()[0]) == ns
    for name, w in want.items():
        a = np.fromfile(tmp_path / "stats" / f"{name}.bin").reshape(w.shape, order="F")
        scale = np.abs(w).max()
        assert np.abs(a - w).max() <= 1e-9 * (scale if scale > 0 else 1.0), name


def test_fortran_host_slab_dumps(tc, oracle, tmp_path):
    """rcuts.in radii -> kslab planes (InitDumpRCuts), and the r-cut, theta-cut,
    streamwise-average and full-field dumps of the last step equal the
    oracle's at the same state; fullff equals the exit fields."""
    if not os.path.exists(HOST):
        pytest.fail("tc_host is not built")
    p = _params(tc)
    p.update(tsta=0.0, nslabr=2, dpthcut=1, dpstave=1, tframe2d=1e-9, dpfullf=1, tframe3d=1e-9)
    for d in ("rcuts", "thcut", "stave", "fullff", "stats"):
        (tmp_path / d).mkdir()
    o = oracle.Oracle(p)
    rm = o.vec("rm")[:p["nrm"]]
    radii = [0.5 * (rm[3] + rm[4]), 0.5 * (rm[10] + rm[11])]   # between planes 4,5 and 11,12
    (tmp_path / "rcuts.in").write_text("POSITION OF SLABS\n" + " ".join("%.17g" % r for r in radii) + "\n")
    tq, f = _run_host(tmp_path, p, fused=True)
    kslab = [int(x) for x in open(tmp_path / "rcuts" / "rcutslocs.out").read().split()]
    assert kslab == [5, 12]
    # the oracle at the same final state
    tq_o, f_o = _oracle_run(tc, oracle, p)
    stamp = "%08d" % int(np.floor(1000.0 * tq[-1, 0] + 0.005 + 0.5))
    o2 = oracle.Oracle(p)
    for n in ("qt", "qr", "qz"):
        o2.set(n, f_o[n])
    nzm, nthm, nrm = p["nzm"], p["nthm"], p["nrm"]
    for m, k in enumerate(kslab, start=1):
        want = o2.dump_rcut(k)
        for q, a in zip(("vt", "vr", "vz"), want):
            got = np.fromfile(tmp_path / "rcuts" / f"slab{m}{q}_{stamp}.bin").reshape((nzm, nthm), order="F")
            assert np.abs(got - a).max() <= 1e-9 * np.abs(a).max(), (m, q)
    for sub, fn, names in (("thcut", o2.dump_thcut, ["slabvt_", "slabvr_", "slabvz_"]),
                           ("stave", o2.dump_streamwise_avg, ["slabvt_ax_", "slabvr_ax_", "slabvz_ax_"])):
        for nm, a in zip(names, fn()):
            got = np.fromfile(tmp_path / sub / f"{nm}{stamp}.bin").reshape((nrm, nzm), order="F")
            assert np.abs(got - a).max() <= 1e-9 * np.abs(a).max(), (sub, nm)
    shape = (nrm + 1, nzm, nthm)
    for n in ("qt", "qr", "qz"):
        got = np.fromfile(tmp_path / "fullff" / f"{n}_{stamp}.bin").reshape(shape, order="F")
        assert np.array_equal(got, f[n]), n


def test_fortran_host_restart_onto_a_finer_grid(tc, oracle, tmp_path):
    """NREAD=1 with continua_master.txt from a coarser run: the host
    interpolates on the device (ReadFlowField + interp.F90), continues the
    time of the old run, and the run equals the same restart driven from
    Python through the C-ABI."""
    if not os.path.exists(HOST):
        pytest.fail("tc_host is not built")
    base = tc.configs.scaled(tc.configs.small_64(), 16, 12, 17)
    base.update(ntst=1, tout=1e-9, sizepert=0.1, omegapert=0.5, idtv=1, dtmin=1e-3, dtmax=5e-3,
                limitcfl=0.5, virotext=-0.2, istr=6, str=2.0)
    a = tmp_path / "coarse"
    a.mkdir()
    _run_host(a, base, fused=True)
    master = open(a / "continua_master.txt").read().split()
    t_old = float(master[3])
    fine = dict(base, nthm=32, nzm=24, nrm=32, nread=1, ireset=0, ntst=3)
    b = tmp_path / "fine"
    b.mkdir()
    for f in ("qt", "qr", "qz"):
        (b / f"continua_{f}.bin").write_bytes((a / f"continua_{f}.bin").read_bytes())
    (b / "continua_master.txt").write_text(open(a / "continua_master.txt").read())
    write_bou(b / "bou.in", fine)
    r = subprocess.run([HOST, "--fused"], cwd=b, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "exit code 555" in r.stdout, r.stdout + r.stderr
    assert "Interpolating new grid" in r.stdout
    tq = np.loadtxt(b / "torquenu_full.out", ndmin=2)
    assert tq[0, 0] > t_old
    # the same restart through the Python mirror
    shape_old = (base["nrm"] + 1, base["nzm"], base["nthm"])
    dev = tc.Solver(fine)
    for f in ("qt", "qr", "qz"):
        old = np.fromfile(a / f"continua_{f}.bin").reshape(shape_old, order="F")
        dev.regrid(f, old, base["istr"], base["str"])
    dev.set_time(t_old, 0.0)
    rows = []
    for n in range(fine["ntst"] + 1):
        code, d = dev.step(n)
        rows.append((d["time"], d["torque_in"], d["torque_out"]))
    assert code == 555
    np.testing.assert_allclose(tq, np.array(rows), rtol=1e-10)
    shape = (fine["nrm"] + 1, fine["nzm"], fine["nthm"])
    for f in ("qt", "qr", "qz"):
        got = np.fromfile(b / f"continua_{f}.bin").reshape(shape, order="F")
        want = dev.get(f)
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), f
