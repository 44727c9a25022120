"""The served sweep's request loop on the CPU (no GPU): admission in arrival
order by free slot, HBM footprint and resident grid points, backfill past a
job that does not fit, and every job ending after its own number of steps.

The device batch is replaced by a fake with the same interface that counts
steps on the host; the host-side set-up (seeded fields, grid vectors, the
fixed dt) is the library's own.
"""
import importlib
import threading
import time

import numpy as np
import pytest


class FakeBatch:
    """tc.Batch stand-in: each advance moves every instance `nsteps` steps
    (a little wall time per step), ending it at NTST with code 555."""

    log = []
    lock = threading.Lock()

    def __init__(self):
        self.inst = {}
        self.next = 0

    def add(self, p, tuning=None):
        i = self.next
        self.next += 1
        self.inst[i] = dict(p=p, steps=0, code=0)
        return i

    def solver(self, i):
        return FakeSolver()

    def advance(self, nsteps, graph=True, sync_every=16):
        time.sleep(0.001 * nsteps)
        for i, s in self.inst.items():
            if s["code"]:
                continue
            s["steps"] = min(s["steps"] + nsteps, s["p"]["ntst"] + 1)
            if s["steps"] == s["p"]["ntst"] + 1:
                s["code"] = 555
        with FakeBatch.lock:
            FakeBatch.log.append(len([s for s in self.inst.values() if not s["code"]]))
        return sum(1 for s in self.inst.values() if not s["code"])

    def status(self, i):
        s = self.inst[i]
        return s["steps"], s["code"], {}

    def remove(self, i):
        del self.inst[i]

    def close(self):
        pass


class FakeSolver:
    def set(self, name, a):
        assert a.ndim == 3

    def project(self):
        pass

    def get(self, name):
        return np.zeros((1, 1, 1))


class FakeTC:
    def __init__(self, tc):
        self._tc = tc
        self.Batch = FakeBatch

    def __getattr__(self, name):  # host-only library calls pass through
        return getattr(self._tc, name)


def _jobs(sw, grids, arrivals, steps):
    rng = np.random.default_rng(3)
    return [(t, sw._instance(rng, g, s), 100 + q)
            for q, (g, t, s) in enumerate(zip(grids, arrivals, steps))]


def test_admission_order_budget_and_backfill(tc):
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    big, small = (48, 48, 33), (16, 16, 9)
    jobs = _jobs(sw, [big, big, small, small], [0.0, 0.01, 0.02, 0.03], [40, 40, 8, 8])
    fb, fs = sw.footprint_bytes(jobs[0][1]), sw.footprint_bytes(jobs[2][1])
    # room for one big job and one small one at a time
    srv = sw.Server(FakeTC(tc), hbm_budget=fb + fs + 1, chunk=4, max_slots=8, workers=2)
    rec = srv.run(jobs)
    for (t, p, _), r in zip(jobs, rec):
        assert r["code"] == 555 and r["steps"] == p["ntst"] + 1, r
        assert t <= r["admitted"] <= r["done"], r
    # the second big job waits for the first to leave; the small jobs are
    # admitted past it (backfill), the second small one after the first left
    assert rec[1]["admitted"] >= rec[0]["done"]
    assert rec[2]["admitted"] < rec[1]["admitted"]
    assert rec[3]["admitted"] >= rec[2]["done"]


def test_slots_and_resident_points_limit_admission(tc):
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    g = (16, 16, 9)
    jobs = _jobs(sw, [g] * 6, [0.0] * 6, [12] * 6)
    # two slots: jobs are admitted in arrival order, two at a time
    rec = sw.Server(FakeTC(tc), chunk=4, max_slots=2, workers=2).run(jobs)
    assert all(r["code"] == 555 for r in rec)
    adm = [r["admitted"] for r in rec]
    assert adm == sorted(adm)
    for q in range(2, 6):
        done_before = sorted(r["done"] for r in rec[:q])
        assert adm[q] >= done_before[q - 2] - 1e-9
    # a resident-point cap of one job's points: one job at a time
    cells = g[0] * g[1] * g[2]
    rec = sw.Server(FakeTC(tc), chunk=4, max_cells=cells, workers=2).run(jobs[:3])
    for q in (1, 2):
        assert rec[q]["admitted"] >= rec[q - 1]["done"]


def test_one_at_a_time_baseline_serves_in_arrival_order(tc):
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")

    class Solver(FakeSolver):
        def __init__(self, p, tuning=None):
            self.p = p

        def run(self, ntime0, nsteps):
            time.sleep(0.0005 * nsteps)
            return 555, self.p["ntst"], {}

        def close(self):
            pass

    ftc = FakeTC(tc)
    ftc.Solver = Solver
    g = (16, 16, 9)
    jobs = _jobs(sw, [g] * 4, [0.0, 0.05, 0.06, 0.3], [20] * 4)
    rec = sw.run_one_at_a_time(ftc, jobs)
    for q, ((t, p, _), r) in enumerate(zip(jobs, rec)):
        assert r["code"] == 555 and r["steps"] == p["ntst"] + 1
        assert r["admitted"] >= t
        if q:
            assert r["admitted"] >= rec[q - 1]["done"]


def test_shortest_first_order_admits_small_jobs_first(tc):
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    big, small = (48, 48, 33), (16, 16, 9)
    jobs = _jobs(sw, [big, big, small, small], [0.0, 0.0, 0.0, 0.0], [40, 40, 8, 8])
    rec = sw.Server(FakeTC(tc), chunk=4, max_slots=1, order="shortest", workers=2).run(jobs)
    adm = [r["admitted"] for r in rec]
    assert max(adm[2], adm[3]) < min(adm[0], adm[1])
    assert adm[0] < adm[1] and adm[2] < adm[3]  # arrival order among equals
    with pytest.raises(ValueError):
        sw.Server(FakeTC(tc), order="random")


def test_admission_order_holds_when_jobs_finish_during_a_pass(tc):
    """A job finishing while the server walks the waiting list must not let a
    later job of the same size pass one that was just turned away: each pass
    decides on one snapshot of the resident set (repeated, since the finish
    has to fall between two admission checks)."""
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    g = (16, 16, 9)
    jobs = _jobs(sw, [g] * 3, [0.0] * 3, [12] * 3)
    cells = g[0] * g[1] * g[2]
    for _ in range(20):
        rec = sw.Server(FakeTC(tc), chunk=4, max_cells=cells, workers=2).run(jobs)
        adm = [r["admitted"] for r in rec]
        assert adm == sorted(adm), rec
        for q in (1, 2):
            assert rec[q]["admitted"] >= rec[q - 1]["done"], rec


def test_batch_autotune_times_each_candidate_on_the_whole_group(tc):
    """batch_autotune builds one batch per candidate with all of a shape's
    instances and keeps the fastest (ties to the lone-instance tuning)."""
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    seen = []

    class TimedBatch(FakeBatch):
        def __enter__(self):
            return self

        def __exit__(self, *a):
            pass

        def add(self, p, tuning=None):
            seen.append((p["nthm"], dict(tuning)))
            i = self.next
            self.next += 1
            self.inst[i] = dict(p=p, steps=0, code=0, tuning=tuning)
            return i

        def advance(self, nsteps, graph=True, sync_every=16):
            # shape 32: lpt 2 with deferral is fastest; shape 16: nothing
            # beats the lone-instance tuning by 2%
            t = 0.0
            for s in self.inst.values():
                tu = s["tuning"]
                if s["p"]["nthm"] == 32:
                    t += 0.004 if (tu["helm_lpt"], tu["defer"]) == (2, 2) else 0.008
                else:
                    t += 0.004
            time.sleep(t * nsteps)
            return len(self.inst)

    ftc = FakeTC(tc)
    ftc.Batch = TimedBatch
    base = {"helm_lpt": 1, "defer": 1, "poisson_pq": 13}
    groups = {(32, 16, 9): [{"nthm": 32}] * 3, (16, 16, 9): [{"nthm": 16}] * 2}
    out = sw.batch_autotune(ftc, groups, steps=2, base={(32, 16, 9): base, (16, 16, 9): base})
    assert out[(32, 16, 9)] == dict(base, helm_lpt=2, defer=2)
    assert out[(16, 16, 9)] == base
    # four candidates per shape, each with the shape's whole group
    assert sum(1 for n, _ in seen if n == 32) == 4 * 3
    assert sum(1 for n, _ in seen if n == 16) == 4 * 2


def test_checkpoints_follow_the_cadence_and_resume_continues(tc, tmp_path):
    """checkpoint_every: a job's state is written at the first chunk boundary
    past every multiple and when it ends (each write replacing the last, the
    end also recorded with its exit code); run(resume=True) continues a job
    from its checkpoint's step, and reports a job with nothing left to run,
    or one that ended by an error code, without running it."""
    import json
    import os
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    writes = []

    class CkptBatch(FakeBatch):
        def write_restart(self, i, path):
            writes.append(self.inst[i]["steps"])
            with open(path, "w") as f:
                json.dump({"steps": self.inst[i]["steps"]}, f)

        def read_restart(self, i, path):
            with open(path) as f:
                n = json.load(f)["steps"]
            self.inst[i]["steps"] = n
            return n

    def restart_info(path):  # the fake checkpoints' header
        with open(path) as f:
            return {"ntime": json.load(f)["steps"]}

    ftc = FakeTC(tc)
    ftc.Batch = CkptBatch
    ftc.restart_info = restart_info
    g = (16, 16, 9)
    jobs = _jobs(sw, [g], [0.0], [20])  # ntst 19: 20 steps
    srv = sw.Server(ftc, chunk=4, workers=1, checkpoint_dir=str(tmp_path), checkpoint_every=6)
    rec = srv.run(jobs)
    assert rec[0]["code"] == 555 and rec[0]["steps"] == 20
    assert writes == [8, 12, 20], writes  # past 6, past 12, the end
    assert sorted(os.listdir(tmp_path)) == ["job0000.end", "job0000.tcr"]
    with open(tmp_path / "job0000.end") as f:
        assert json.load(f) == {"steps": 20, "code": 555}
    # resumed as it was: nothing left to run, reported without stepping
    rec = srv.run(jobs, resume=True)
    assert (rec[0]["steps"], rec[0]["code"], rec[0]["resumed_at"]) == (20, 555, 20)
    assert writes == [8, 12, 20], writes
    # the same job extended to 40 steps continues from step 20
    longer = [(t, dict(p, ntst=39), s) for t, p, s in jobs]
    rec = srv.run(longer, resume=True)
    assert rec[0]["resumed_at"] == 20 and rec[0]["steps"] == 40 and rec[0]["code"] == 555
    assert writes[-1] == 40
    # a job that ended by an error code is not run again, even extended
    with open(tmp_path / "job0000.end", "w") as f:
        json.dump({"steps": 40, "code": 165}, f)
    rec = srv.run([(t, dict(p, ntst=59), s) for t, p, s in jobs], resume=True)
    assert (rec[0]["steps"], rec[0]["code"], rec[0]["resumed_at"]) == (40, 165, 40)
    assert writes[-1] == 40
    with pytest.raises(ValueError):
        sw.Server(ftc, checkpoint_every=5)
