"""The batched-sweep and served-sweep workloads at their own grid shapes.

- Batched sweep (48 uneven instances over 64^2x65, 128^2x97, 192x128x129,
  256^2x193): one instance of each shape, with the sweep's own parameter
  draws and fixed dt, advanced together in one batch (grouped launches,
  hipGraph replay).  Each instance ends bit for bit where the plain
  single-instance path ends, and the two smaller shapes also match the
  oracle (relative L2 <= 1e-10 on qt, qr, qz and pressure).
- Served sweep (grids up to 384^2x289 with an oscillating inner cylinder):
  a reduced job list that includes the largest shape runs through the
  request loop; every job's final fields equal the same job run alone on the
  plain path, bit for bit, and every job ends with code 555 after its own
  number of steps.
"""
import importlib

import numpy as np
import pytest

from helpers import compare_fields

pytestmark = pytest.mark.gpu

FIELDS = ("qt", "qr", "qz", "pr")


def _sweep():
    return importlib.import_module("taylor-couette-afid_amd.sweep")


def _one_per_shape(sw, steps):
    seen, out = set(), []
    for p, seed in sw.sweep_jobs(steps=steps):
        g = (p["nthm"], p["nzm"], p["nrm"])
        if g not in seen:
            seen.add(g)
            out.append((p, seed))
    assert sorted(seen) == sorted(sw.SWEEP_GRIDS)
    return out


def test_sweep_shapes_batched_match_plain_and_oracle(tc, oracle):
    sw = _sweep()
    nsteps = 6
    jobs = _one_per_shape(sw, nsteps)
    starts = []
    for p, seed in jobs:
        f = sw.start_fields(tc, p, seed)
        s = tc.Solver(p)
        sw.load_and_project(s, f)
        starts.append({n: s.get(n) for n in ("qt", "qr", "qz")})
        s.close()
    print("sweep shapes in one batch:", [(p["nthm"], p["nzm"], p["nrm"], round(p["dtmax"], 6))
                                         for p, _ in jobs], flush=True)
    with tc.Batch() as b:
        ids = []
        for (p, _), st in zip(jobs, starts):
            i = b.add(p)
            s = b.solver(i)
            for n, a in st.items():
                s.set(n, a)
            ids.append(i)
        assert b.groups()[0] == 4
        # every instance ends at NTST (code 555): none is left active
        assert b.advance(nsteps, graph=True, sync_every=3) == 0
        got = [{f: b.solver(i).get(f) for f in FIELDS} for i in ids]
        status = [b.status(i) for i in ids]
    for (p, _), st, g, (steps, code, _) in zip(jobs, starts, got, status):
        assert (steps, code) == (nsteps, 555), (p["nthm"], steps, code)
        s = tc.Solver(p)
        for n, a in st.items():
            s.set(n, a)
        for n in range(nsteps):
            c, _ = s.step(n)
        assert c == 555
        for f in FIELDS:
            a = s.get(f)
            assert np.array_equal(g[f], a), (p["nthm"], f, np.abs(g[f] - a).max())
        if p["nthm"] * p["nzm"] * p["nrm"] <= 128 * 128 * 97:
            orc = oracle.Oracle(p)
            for n, a in st.items():
                orc.set(n, a)
            for n in range(nsteps):
                c2, _ = orc.step(n)
            assert c2 == 555
            e = compare_fields(s, orc, p["nrm"])
            print(f"  {p['nthm']}x{p['nzm']}x{p['nrm']} vs oracle {e}", flush=True)
            orc.close()
        s.close()


def test_served_jobs_up_to_384_same_alone_and_in_the_mix(tc):
    sw = _sweep()
    rng = np.random.default_rng(5005)
    grids = [(64, 64, 65), (384, 384, 289), (128, 128, 97), (64, 64, 65), (256, 256, 193)]
    jobs = []
    for q, g in enumerate(grids):
        p = sw._instance(rng, g, 10 + 2 * q)
        p.update(sizepert=rng.uniform(0.05, 0.2), omegapert=rng.uniform(0.01, 0.1))
        jobs.append((0.02 * q, p, 7000 + q))
    srv = sw.Server(tc, chunk=4)
    rec = srv.run(jobs, keep_fields=True)
    for (t, p0, seed), r in zip(jobs, rec):
        nsteps = p0["ntst"] + 1
        assert r["code"] == 555 and r["steps"] == nsteps, r
        p = dict(p0)
        f = sw.start_fields(tc, p, seed)
        s = tc.Solver(p)
        sw.load_and_project(s, f)
        for n in range(nsteps):
            code, _ = s.step(n)
        assert code == 555
        for name, got in zip(FIELDS, r["fields"]):
            want = s.get(name)
            assert np.array_equal(got, want), (p["nthm"], name, np.abs(got - want).max())
        s.close()
    print("served:", [(p["nthm"], round(r["done"] - r["arrival"], 2)) for (_, p, _), r in zip(jobs, rec)],
          flush=True)
