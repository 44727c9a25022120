"""Served sweep: the continuous-batching request loop.

Jobs arrive while others run, join the resident batch at step boundaries and
leave when done; each job's final fields equal those of the same job run
alone on the plain single-instance path (bit for bit), and every job ends
with code 555 after its own number of steps.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(sw, steps):
    rng = np.random.default_rng(77)
    grids = [(32, 24, 33), (48, 32, 25), (32, 24, 33), (24, 40, 17), (48, 32, 25)]
    out = []
    for q, g in enumerate(grids):
        p = sw._instance(rng, g, steps + 3 * q)
        p.update(sizepert=0.1, omegapert=0.05)
        out.append((0.05 * q, p, 900 + q))
    return out


def test_served_results_same_alone_and_in_the_mix(tc):
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    jobs = _jobs(sw, 12)
    srv = sw.Server(tc, chunk=4)
    rec = srv.run(jobs, keep_fields=True)
    for (t, p0, seed), r in zip(jobs, rec):
        nsteps = p0["ntst"] + 1
        assert r["code"] == 555 and r["steps"] == nsteps, r
        assert r["done"] >= r["admitted"] >= t
        p = dict(p0)
        f = sw.start_fields(tc, p, seed)
        s = tc.Solver(p)
        sw.load_and_project(s, f)
        for n in range(nsteps):
            code, _ = s.step(n)
        assert code == 555
        for name, got in zip(("qt", "qr", "qz", "pr"), r["fields"]):
            want = s.get(name)
            assert np.array_equal(got, want), (seed, name, np.abs(got - want).max())
        s.close()


def test_served_jobs_resume_from_their_checkpoints(tc, tmp_path):
    """Each job checkpoints through the restart format as it runs; a second
    server run that extends the jobs continues them from their checkpoints,
    and ends with the same bits as one uninterrupted run."""
    sw = importlib.import_module("taylor-couette-afid_amd.sweep")
    short = _jobs(sw, 10)[:3]
    long = [(t, dict(p, ntst=p["ntst"] + 8), s) for t, p, s in short]
    srv = sw.Server(tc, chunk=4, checkpoint_dir=str(tmp_path), checkpoint_every=4)
    first = srv.run(short)
    assert all(r["code"] == 555 for r in first)
    resumed = srv.run(long, keep_fields=True, resume=True)
    straight = sw.Server(tc, chunk=4).run(long, keep_fields=True)
    for (t, p, s), r0, r1, r2 in zip(long, first, resumed, straight):
        assert r1["resumed_at"] == r0["steps"] and "resumed_at" not in r2
        assert r1["code"] == r2["code"] == 555 and r1["steps"] == r2["steps"] == p["ntst"] + 1
        for name, a, b in zip(("qt", "qr", "qz", "pr"), r1["fields"], r2["fields"]):
            assert np.array_equal(a, b), (s, name, np.abs(a - b).max())
