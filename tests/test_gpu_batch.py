"""Batched sweep: grouped launches over many instances.

- every instance of a batch follows exactly the single-instance fused path
  (tc_step): fields bit-identical after several steps, with the inner
  cylinder oscillating;
- hipGraph replay gives the same fields as eager launches, replay after
  replay, with host synchronisation between replays;
- an instance gives the same result alone and in a mix of other shapes;
- the exit codes and their steps are those of the main.F90 loop;
- the autotuner returns a schedule that runs and matches the default one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("qt", "qr", "qz", "pr")


def _params(tc, nthm, nzm, nrm, **kw):
    p = tc.configs.scaled(tc.configs.small_64(), nthm, nzm, nrm)
    p.update(idtv=0, dtmax=2e-3, dtmin=2e-3, limitcfl=1e6, tout=1e30, tmax=1e30,
             sizepert=0.1, omegapert=0.5, ntst=1000)
    p.update(kw)
    return p


def _start(tc, p, seed):
    """Projected perturbed Couette start, computed once on a single context."""
    s = tc.Solver(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, 1e-2)
    s.set("qt", qt); s.set("qr", qr); s.set("qz", qz)
    s.project()
    st = {f: s.get(f) for f in ("qt", "qr", "qz")}
    s.close()
    return st


def _load(solver, st):
    for f, a in st.items():
        solver.set(f, a)


def _mix(tc):
    return [
        (_params(tc, 32, 24, 33, ren=800.0), 11),
        (_params(tc, 32, 24, 33, ren=1500.0, rext=1.6, virotext=-0.2), 12),
        (_params(tc, 48, 32, 25, ren=1000.0, lamb=2.0), 13),
        (_params(tc, 24, 40, 17, ren=600.0, rext=1.25), 14),
    ]


def _single(tc, p, st, nsteps, tuning=None):
    s = tc.Solver(p, tuning=tuning)
    _load(s, st)
    diags = []
    for n in range(nsteps):
        code, d = s.step(n)
        diags.append((code, d))
    out = {f: s.get(f) for f in FIELDS}
    s.close()
    return out, diags


# default (off at these sizes), forced by the environment, forced per
# instance by the tuning
@pytest.mark.parametrize("defer", ["1", "2", "tuning"])
def test_batch_matches_single_instance_bitwise(tc, monkeypatch, defer):
    tuning = None
    if defer == "tuning":
        tuning = {"defer": 2}
    else:
        monkeypatch.setenv("TC_DEFER_CORRECTION", defer)
    nsteps = 5
    mix = _mix(tc)
    starts = [_start(tc, p, seed) for p, seed in mix]
    with tc.Batch() as b:
        ids = []
        for (p, _), st in zip(mix, starts):
            i = b.add(p, tuning)
            _load(b.solver(i), st)
            ids.append(i)
        ng, launches = b.groups()
        assert ng == 3, ng
        assert b.advance(nsteps, graph=True, sync_every=2) == len(mix)
        for (p, _), st, i in zip(mix, starts, ids):
            assert b.solver(i).tuning()["defer"] == (2 if defer != "1" else 1)
            ref, diags = _single(tc, p, st, nsteps, tuning)
            got = b.solver(i)
            for f in FIELDS:
                a = got.get(f)
                assert np.array_equal(a, ref[f]), (i, f, np.abs(a - ref[f]).max())
            steps, code, d = b.status(i)
            assert steps == nsteps and code == 0
            # the last logging step is ntime = 1; maxima are the same bits,
            # the torque sums run in another order where the single instance
            # evaluates them inside its last correction pass
            d1 = diags[1][1]
            for k in ("dmax", "vmax1", "vmax2", "vmax3"):
                assert d[k] == d1[k], (k, d[k], d1[k])
            for k in ("torque_in", "torque_out"):
                assert abs(d[k] / d1[k] - 1) < 1e-13, (k, d[k], d1[k])


def test_odd_shapes_in_a_batch_match_single_instance(tc):
    """Odd theta or z sizes in one batch (the grouped kernels' paths for odd
    nthm and theta-FFT row pairs that straddle r-planes), two instances of
    one odd shape in a group: bitwise the single-instance runs."""
    mix = [(_params(tc, 27, 20, 13, ren=700.0), 31), (_params(tc, 27, 20, 13, ren=900.0), 32),
           (_params(tc, 40, 45, 17, ren=800.0, rext=1.5), 33),
           (_params(tc, 45, 27, 9, ren=600.0), 34)]
    starts = [_start(tc, p, seed) for p, seed in mix]
    nsteps = 4
    with tc.Batch() as b:
        ids = []
        for (p, _), st in zip(mix, starts):
            i = b.add(p)
            _load(b.solver(i), st)
            ids.append(i)
        assert b.groups()[0] == 3
        assert b.advance(nsteps, graph=True, sync_every=2) == len(mix)
        for (p, _), st, i in zip(mix, starts, ids):
            ref, _ = _single(tc, p, st, nsteps)
            assert b.status(i)[:2] == (nsteps, 0)
            for f in FIELDS:
                a = b.solver(i).get(f)
                assert np.array_equal(a, ref[f]), (i, f, np.abs(a - ref[f]).max())


def test_batch_graph_replays_match_eager(tc):
    mix = _mix(tc)[:3]
    starts = [_start(tc, p, seed) for p, seed in mix]
    results = []
    for graph in (False, True):
        with tc.Batch() as b:
            ids = []
            for (p, _), st in zip(mix, starts):
                i = b.add(p)
                _load(b.solver(i), st)
                ids.append(i)
            snaps = []
            for rep in range(3):  # three advances: the graph is replayed, host syncs between
                b.advance(2, graph=graph, sync_every=1)
                snaps.append([b.solver(i).get("qt") for i in ids])
            results.append(snaps)
    for rep in range(3):
        for q in range(len(mix)):
            assert np.array_equal(results[0][rep][q], results[1][rep][q]), (rep, q)


def test_same_alone_and_in_the_mix(tc):
    mix = _mix(tc)
    p0, seed0 = mix[1]
    st0 = _start(tc, p0, seed0)
    out = []
    for others in (False, True):
        with tc.Batch() as b:
            i0 = b.add(p0)
            _load(b.solver(i0), st0)
            if others:
                for p, seed in mix[:1] + mix[2:]:
                    i = b.add(p)
                    _load(b.solver(i), _start(tc, p, seed))
            b.advance(4, graph=True, sync_every=4)
            out.append({f: b.solver(i0).get(f) for f in FIELDS})
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[1][f]), f


def test_batch_exit_codes_match_main_loop(tc):
    """The exit code and step of each instance are those of the main.F90
    loop, and an instance that fails on the device (165 at a step start,
    169 in a logging pass) takes no step after the failing one even though
    the host reads the results only every sync_every steps: its fields, time
    and diagnostics equal the single-instance loop's at the failing step."""
    base = _params(tc, 32, 24, 33)
    cases = [dict(ntst=3), dict(tmax=0.0045), dict(limitcfl=1e-6), dict(ntst=6, resid=1e-30)]
    st = _start(tc, base, 21)
    with tc.Batch() as b:
        ids = []
        for kw in cases:
            p = dict(base, **kw)
            i = b.add(p)
            _load(b.solver(i), st)
            ids.append((i, p))
        b.advance(10, graph=True, sync_every=3)
        for i, p in ids:
            steps, code, out9 = b.status(i)
            got = {f: b.solver(i).get(f) for f in FIELDS}
            s = tc.Solver(p)
            _load(s, st)
            want = None
            for n in range(10):
                c, d = s.step(n)
                if c:
                    want = (n + 1, c)
                    break
            assert (steps, code) == want, (p, steps, code, want)
            for f in FIELDS:
                assert np.array_equal(got[f], s.get(f)), (p, f)
            # cfl: the batch's step-start CFL pass and the single instance's
            # cached logging CFL evaluate the same expression
            assert abs(out9["cfl"] / d["cfl"] - 1) <= 1e-14, (p, out9["cfl"], d["cfl"])
            logged = n == 1
            for k in tc.Batch.STATUS[1:3] + (tc.Batch.STATUS[3:] if logged else []):
                if k.startswith("torque"):
                    assert abs(out9[k] / d[k] - 1) < 1e-13, (p, k, out9[k], d[k])
                else:
                    assert out9[k] == d[k], (p, k, out9[k], d[k])
            s.close()


def test_fixed_dt_logging_cfl_is_scaled_by_dt(tc, oracle):
    """main.F90:172,201: in fixed-dt mode both the step-start CFL and the
    logging pass's CFL are reported times dt."""
    p = _params(tc, 32, 24, 33, tout=1e30)
    st = _start(tc, p, 5)
    s = tc.Solver(p)
    _load(s, st)
    orc = oracle.Oracle(p)
    for f, a in st.items():
        orc.set(f, a)
    for n in range(3):
        c, d = s.step(n)
        c2, d2 = orc.step(n)
        assert abs(d["cfl"] / d2["cfl"] - 1) < 1e-12, (n, d["cfl"], d2["cfl"])
        if n == 1:  # the logging step
            cfl_raw = s.CalcMaxCFL()
            assert abs(d["cfl"] - cfl_raw * p["dtmax"]) <= 1e-15 * cfl_raw * p["dtmax"]
    s.close()


def test_autotune_schedule_runs_and_matches_default(tc):
    p = _params(tc, 64, 48, 33)
    t = tc.autotune(p, steps=2)
    assert t["helm_lpt"] in (1, 2) and t["poisson_pq"] in (13, 26)
    assert t["fft_z_nt"] in (256, 512) and t["fft_pad"] in (1, -1)
    st = _start(tc, p, 31)
    a = tc.Solver(p, tuning=t)
    b = tc.Solver(p)
    _load(a, st)
    _load(b, st)
    for n in range(3):
        a.step(n)
        b.step(n)
    for f in ("qt", "qr", "qz"):
        x, y = a.get(f), b.get(f)
        assert np.linalg.norm(x - y) <= 1e-12 * np.linalg.norm(y), f
    assert a.tuning()["poisson_pq"] == t["poisson_pq"]
    a.close()
    b.close()


def test_tuning_values_outside_the_documented_sets_are_rejected(tc):
    p = _params(tc, 32, 24, 17)
    for bad in ({"poisson_pq": 7}, {"helm_lpt": 3}, {"defer": 3}):
        with pytest.raises(tc.TCError):
            tc.Solver(p, tuning=bad)
    s = tc.Solver(p)
    with pytest.raises(tc.TCError):
        s.set_tuning({"poisson_pq": 20})
    assert s.tuning()["poisson_pq"] == 13
    s.close()


def test_batch_checkpoint_continues_bit_for_bit(tc, tmp_path):
    """tc_batch_write_restart mid-run, tc_batch_read_restart into a fresh
    batch (and tc_read_restart on the single-instance path): both continue
    as the uninterrupted batch does, bit for bit."""
    mix = _mix(tc)[2:4]
    starts = [_start(tc, p, seed) for p, seed in mix]
    paths = [str(tmp_path / f"inst{q}.tcr") for q in range(len(mix))]
    with tc.Batch() as a:
        ids = []
        for (p, _), st in zip(mix, starts):
            i = a.add(p)
            _load(a.solver(i), st)
            ids.append(i)
        a.advance(6, graph=True, sync_every=4)
        for i, path in zip(ids, paths):
            a.write_restart(i, path)
        a.advance(5, graph=True, sync_every=4)
        want = [{f: a.solver(i).get(f) for f in FIELDS} for i in ids]
        assert all(a.status(i)[:2] == (11, 0) for i in ids)
    with tc.Batch() as b:
        ids = [b.add(p) for p, _ in mix]
        assert [b.read_restart(i, path) for i, path in zip(ids, paths)] == [6, 6]
        with pytest.raises(tc.TCError):
            b.read_restart(ids[0], paths[0])  # already past step 0
        extra = b.add(mix[0][0])
        with pytest.raises(tc.TCError):
            b.read_restart(extra, paths[1])  # another grid
        b.remove(extra)
        # another fixed dt: refused from the header, before any field moves
        p2 = dict(mix[0][0])
        p2.update(dtmax=1e-3, dtmin=1e-3)
        extra = b.add(p2)
        _load(b.solver(extra), starts[0])
        with pytest.raises(tc.TCError, match="DTMAX"):
            b.read_restart(extra, paths[0])
        for f, v in starts[0].items():
            assert np.array_equal(b.solver(extra).get(f), v), f
        b.remove(extra)
        # a checkpoint past the instance's NTST: nothing left to run
        extra = b.add(dict(mix[0][0], ntst=5))
        with pytest.raises(tc.TCError, match="past NTST"):
            b.read_restart(extra, paths[0])
        b.remove(extra)
        info = tc.restart_info(paths[0])
        assert (info["nthm"], info["nzm"], info["nrm"], info["ntime"]) == (48, 32, 25, 6)
        assert info["dt"] == 2e-3 and abs(info["time"] - 6 * 2e-3) < 1e-15
        b.advance(5, graph=True, sync_every=2)
        for i, w in zip(ids, want):
            assert b.status(i)[:2] == (11, 0)
            for f in FIELDS:
                assert np.array_equal(b.solver(i).get(f), w[f]), (i, f)
    for (p, _), path, w in zip(mix, paths, want):
        s = tc.Solver(p)
        assert s.read_restart(path) == 6
        code, last, _ = s.run(6, 5)
        assert (code, last) == (0, 10)
        for f in FIELDS:
            assert np.array_equal(s.get(f), w[f]), f
        s.close()


def test_batch_autotune_returns_schedules_that_run(tc):
    """sweep.batch_autotune on two small shape groups: a valid schedule per
    shape, built on each shape's lone-instance tuning, and a batch with those
    schedules steps like the single-instance path with the same schedule."""
    import importlib
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    mix = _mix(tc)
    groups = {}
    for p, _ in mix:
        groups.setdefault((p["nthm"], p["nzm"], p["nrm"]), []).append(p)
    tun = sw.batch_autotune(tc, groups, steps=2)
    assert set(tun) == set(groups)
    for sh, t in tun.items():
        assert t["helm_lpt"] in (1, 2) and t["defer"] in (1, 2, -1), (sh, t)
    p, seed = mix[0]
    st = _start(tc, p, seed)
    t = tun[(p["nthm"], p["nzm"], p["nrm"])]
    with tc.Batch() as b:
        i = b.add(p, t)
        _load(b.solver(i), st)
        b.advance(3, graph=True, sync_every=3)
        got = {f: b.solver(i).get(f) for f in FIELDS}
    ref, _ = _single(tc, p, st, 3, t)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
