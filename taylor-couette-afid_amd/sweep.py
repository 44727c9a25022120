"""Parameter sweeps of many Taylor-Couette instances on one device.

Two workloads:

- ``sweep_jobs``: the batched sweep, 48 independent instances of uneven size
  (grids from {64^2x65, 128^2x97, 192x128x129, 256^2x193}), eta uniform in
  [0.5, 0.909], 1/NU log-uniform in [1e3, 1e4], VIROTEXT/VIROTINT uniform in
  [-0.5, 0.2], LAMB in {2, 4, 8}, each from a seeded 1e-2 perturbation of its
  own Couette profile, fixed dt from a CFL = 0.8 estimate (capped where the
  reference's pressure update would be unstable, see ``pressure_stable_dt``),
  300 RK3 steps.
- ``served_jobs``: the served sweep, Poisson arrivals (mean 2 jobs/s) of the
  same distribution over grids 64^2x65 .. 384^2x289, plus an inner-cylinder
  oscillation of amplitude in [0, 0.2] and frequency in [0.01, 0.1], 1000 RK3
  steps each.

Values the workload descriptions leave open: RINT = 1 (so REXT = 1/eta),
ALX3D = 2 pi, VIROTINT = 1, STR3 = 1.5 with ISTR = 4, RESID = 1e-6.

``Server`` is the request loop of the served form.  Arrived jobs wait in
arrival order; a job is admitted when a slot is free, its device footprint
fits the HBM budget and the resident grid points stay under ``max_cells``
(enough to keep the device saturated; beyond it every resident job would
only slow the others down).  A job that does not fit may be overtaken by a
later one that does (backfill).  Admitted jobs stay resident in one
``Batch`` per grid shape, each advanced by its own thread on its own
streams, so small jobs do not wait for the step of a large one; each
batch's step (all its grouped launches) is replayed as one hipGraph; jobs
join and leave a batch at step boundaries, every ``chunk`` steps.  Host fields
are prepared on worker threads while the device computes, and each job's
time to result (arrival to final step) is recorded.  With a checkpoint
directory, every job's whole state is written as it runs (the restart
format of tc_write_restart) and a later run continues jobs from it.

``batch_autotune`` picks a schedule per grid shape for a batch (timed on
the whole group of that shape); ``run_one_at_a_time`` is the baseline that
serves the same trace one job at a time on the single-instance path.
"""
import concurrent.futures as cf
import json
import math
import os
import threading
import time

import numpy as np

from . import configs

SWEEP_GRIDS = [(64, 64, 65), (128, 128, 97), (192, 128, 129), (256, 256, 193)]
SERVED_GRIDS = SWEEP_GRIDS + [(384, 384, 289)]


def _instance(rng, grid, steps):
    nthm, nzm, nrm = grid
    eta = rng.uniform(0.5, 0.909)
    ren = 10.0 ** rng.uniform(3.0, 4.0)
    ratio = rng.uniform(-0.5, 0.2)
    lamb = float(rng.choice([2.0, 4.0, 8.0]))
    p = dict(configs.BASE)
    p.update(nthm=nthm, nzm=nzm, nrm=nrm, idtv=0, ntst=steps - 1, alx3=2.0 * math.pi,
             rint=1.0, rext=1.0 / eta, lamb=lamb, virotint=1.0, virotext=ratio,
             ren=ren, limitcfl=1.0e30)
    return p


def sweep_jobs(n=48, seed=4004, steps=300):
    """(params, field seed) of the batched sweep; the fixed dt is set by
    ``start_fields``."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        grid = SWEEP_GRIDS[int(rng.integers(len(SWEEP_GRIDS)))]
        out.append((_instance(rng, grid, steps), seed * 1000 + q))
    return out


def served_jobs(n, seed=5005, rate=2.0, steps=1000):
    """(arrival time in s, params, field seed) of the served sweep."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n):
        t += rng.exponential(1.0 / rate)
        grid = SERVED_GRIDS[int(rng.integers(len(SERVED_GRIDS)))]
        p = _instance(rng, grid, steps)
        p.update(sizepert=rng.uniform(0.0, 0.2), omegapert=rng.uniform(0.01, 0.1))
        out.append((t, p, seed * 1000 + q))
    return out


def cfl_estimate(tc, p, qt, qr, qz):
    """CalcMaxCFL of a host field (Fortran layout (nr, nzm, nthm))."""
    nrm = p["nrm"]
    usrm = tc.grid_vector(p, "usrm")[:nrm]
    udx2rm = tc.grid_vector(p, "udx2rm")[:nrm]
    dxt = p["nthm"] * p["lamb"] / (2.0 * math.pi)
    dxz = p["nzm"] / p["alx3"]
    t = qt[:nrm]
    a = np.abs((np.roll(t, -1, axis=2) + t) * 0.5 * dxt * usrm[:, None, None])
    a += np.abs((qr[1:nrm + 1] + qr[:nrm]) * 0.5 * udx2rm[:, None, None])
    z = qz[:nrm]
    a += np.abs((np.roll(z, -1, axis=1) + z) * 0.5 * dxz)
    return max(1e-8, float(a.max()))


def pressure_step_amplification(tc, p, dt):
    """Largest per-step growth factor of a pressure error mode under the
    reference's RK3 step.

    CorrectPressure adds phi - beta L phi with beta = dt/(2 Re) (not
    al*beta), and the implicit part of each substep is radial only.  A
    pressure error mode with radial eigenvalue part x/beta and theta-z part
    y/beta is multiplied in substep s by
    -((1 - al_s) x + y) / (1 + al_s x); the step multiplies the three.  The
    bound takes y at its largest (theta and z Nyquist at the inner radius)
    and x over [0, Gershgorin bound of the radial operator]."""
    nrm = p["nrm"]
    acph = tc.grid_vector(p, "acph")[:nrm]
    usrm = tc.grid_vector(p, "usrm")[:nrm]
    dxt = p["nthm"] * p["lamb"] / (2.0 * math.pi)
    dxz = p["nzm"] / p["alx3"]
    beta = dt / (2.0 * p["ren"])
    y = beta * (4.0 * dxt * dxt * float(np.max(usrm * usrm)) + 4.0 * dxz * dxz)
    x = np.linspace(0.0, beta * 2.0 * float(np.max(np.abs(acph))), 2001)
    g = np.ones_like(x)
    for al in (8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0):
        g = g * ((1.0 - al) * x + y) / (1.0 + al * x)
    return float(g.max())


def pressure_stable_dt(tc, p, dt_hi, safety=0.9):
    """dt_hi if the pressure update is stable there, else `safety` times the
    largest dt below it with a step amplification of at most 1."""
    if pressure_step_amplification(tc, p, dt_hi) <= 1.0:
        return dt_hi
    lo, hi = 0.0, dt_hi
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if pressure_step_amplification(tc, p, mid) <= 1.0:
            lo = mid
        else:
            hi = mid
    return safety * lo


def start_fields(tc, p, seed, amp=1e-2, cfl=0.8):
    """Seeded perturbed Couette start and its fixed dt: the CFL = 0.8
    estimate of the start, reduced by ``pressure_stable_dt`` where the
    reference's pressure update would amplify errors; sets
    p['dtmax'] = p['dtmin'] = dt in place."""
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    dt = pressure_stable_dt(tc, p, cfl / cfl_estimate(tc, p, qt, qr, qz))
    p["dtmax"] = p["dtmin"] = dt
    return qt, qr, qz


def footprint_bytes(p, tuning=None):
    """Device bytes of one instance: eleven fields of nr planes (twelve where
    the RK3 step defers its pressure correction: even nthm and the tuning's
    ``defer`` 2, or 1 (the default, unless TC_DEFER_CORRECTION says 2 or 0)
    with at least 8 Mi grid points), the spectrum, and the small
    per-instance tables."""
    nthm, nzm, nrm = p["nthm"], p["nzm"], p["nrm"]
    sstr = ((nthm // 2 + 1) + 7) // 8 * 8
    mode = int((tuning or {}).get("defer", 0))
    if mode == 0:
        mode = {"0": -1, "2": 2}.get(os.environ.get("TC_DEFER_CORRECTION", "1"), 1)
    defer = (nthm % 2 == 0 and p.get("nsst", 3) == 3
             and (mode == 2 or (mode == 1 and nthm * nzm * nrm >= 8 << 20)))
    nf = 12 if defer else 11
    return 8 * nf * (nrm + 1) * nzm * nthm + 16 * nrm * nzm * sstr + (8 << 20)


def batch_autotune(tc, groups, steps=3, base=None):
    """Per-shape tuning for a Batch.  ``tc.autotune`` times one instance
    alone; in a batch the grouped launches see all instances of a shape at
    once, so both the lines per thread of the Helmholtz kernel (resident
    blocks against tiles) and the deferred correction (which a lone small
    grid, living in the caches, does not repay) can come out the other way.
    For each shape, a batch of that shape's instances (``groups``: shape ->
    list of params; zero fields, the real count) is timed over ``steps``
    graph-replayed steps for each candidate: the lone-instance tuning
    (``base``, else ``tc.autotune``) with helm_lpt 1 and 2, each with the
    correction deferred by size and always.  Returns shape -> tuning."""
    out = {}
    for sh, plist in groups.items():
        t0 = dict((base or {}).get(sh) or tc.autotune(plist[0], steps=steps))
        cands = []
        for lpt in (1, 2):
            for defer in (t0.get("defer", 1), 2):
                c = dict(t0, helm_lpt=lpt, defer=defer)
                if c not in cands:
                    cands.append(c)
        best, best_ms = t0, None
        for c in cands:
            with tc.Batch() as b:
                for p in plist:
                    b.add(p, c)
                b.advance(1, graph=True, sync_every=1)
                t = time.perf_counter()
                b.advance(steps, graph=True, sync_every=steps)
                ms = (time.perf_counter() - t) / steps * 1e3
            if best_ms is None or ms < 0.98 * best_ms:
                best, best_ms = c, ms
        out[sh] = best
    return out


def load_and_project(solver, fields):
    for name, a in zip(("qt", "qr", "qz"), fields):
        solver.set(name, a)
    solver.project()


class _ClassWorker:
    """One Batch of one grid shape, advanced by its own thread: jobs of a
    small shape are not held to the step pace of a large one (their kernels
    interleave with the large shape's on the device).  New jobs are handed
    over through a queue and join the batch at the next chunk boundary."""

    def __init__(self, server, shape):
        self.srv = server
        self.shape = shape
        self.b = server.tc.Batch()
        self.qlock = threading.Lock()
        self.pending = []          # (job index, params, fields, tuning, checkpoint or None)
        self.active = {}           # batch id -> job index
        self.saved = {}            # batch id -> step of its last checkpoint
        self.wake = threading.Event()
        self.stop = False
        self.error = None
        self.thread = threading.Thread(target=self._loop, daemon=True)
        self.thread.start()

    def admit(self, q, p, fields, tuning, resume=None):
        with self.qlock:
            self.pending.append((q, p, fields, tuning, resume))
        self.wake.set()

    def _loop(self):
        srv, b = self.srv, self.b
        try:
            while True:
                if not self.active:
                    self.wake.wait(0.05)
                self.wake.clear()
                if self.stop:
                    break
                with self.qlock:
                    new, self.pending = self.pending, []
                for q, p, fields, tuning, resume in new:
                    iid = b.add(p, tuning)
                    if resume:
                        self.saved[iid] = b.read_restart(iid, resume)
                        srv._resumed(q, self.saved[iid])
                    else:
                        load_and_project(b.solver(iid), fields)
                        self.saved[iid] = 0
                    self.active[iid] = q
                    srv._admitted(q)
                del new
                if not self.active:
                    continue
                b.advance(srv.chunk, graph=srv.graph, sync_every=srv.chunk)
                for iid in list(self.active):
                    steps, code, _ = b.status(iid)
                    every = srv.checkpoint_every
                    if every and (code or steps // every > self.saved[iid] // every):
                        srv._checkpoint(b, iid, self.active[iid], end=(steps, code) if code else None)
                        self.saved[iid] = steps
                    if code:
                        q = self.active.pop(iid)
                        fields = None
                        if srv.keep_fields:
                            s = b.solver(iid)
                            fields = tuple(s.get(f) for f in ("qt", "qr", "qz", "pr"))
                        b.remove(iid)
                        del self.saved[iid]
                        srv._finished(q, steps, code, fields)
        except Exception as e:  # reported by Server.run
            self.error = e
            srv._finished(None, 0, 0, None)

    def close(self):
        self.stop = True
        self.wake.set()
        self.thread.join()
        self.b.close()


def run_one_at_a_time(tc, jobs, tuning=None, ahead=2):
    """The baseline of the served sweep: the same arrival trace served one
    job at a time in arrival order, each job alone on the plain
    single-instance path (tc_run_steps from its start to NTST).  Host fields
    are prepared on a worker thread ahead of time, as the server does.
    Returns one record per job: arrival, admitted, done (s since start),
    steps, code."""
    tuning = tuning or {}
    order = sorted(range(len(jobs)), key=lambda q: jobs[q][0])
    rec = [None] * len(jobs)

    def prepare(q):
        p = dict(jobs[q][1])
        return p, start_fields(tc, p, jobs[q][2])

    with cf.ThreadPoolExecutor(max_workers=max(1, ahead)) as pool:
        futs = {q: pool.submit(prepare, q) for q in order[:ahead]}
        t0 = time.perf_counter()
        for n, q in enumerate(order):
            for q2 in order[n + 1:n + 1 + ahead]:
                if q2 not in futs:
                    futs[q2] = pool.submit(prepare, q2)
            wait = jobs[q][0] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            admitted = time.perf_counter() - t0
            p, fields = futs.pop(q).result()
            sh = (p["nthm"], p["nzm"], p["nrm"])
            s = tc.Solver(p, tuning=tuning.get(sh))
            load_and_project(s, fields)
            del fields
            code, last, _ = s.run(0, p["ntst"] + 1)
            s.close()
            rec[q] = dict(arrival=jobs[q][0], admitted=admitted, done=time.perf_counter() - t0,
                          steps=last + 1, code=code, cells=p["nthm"] * p["nzm"] * p["nrm"])
    return rec


def _work(p):
    """Grid points times steps of one job."""
    return p["nthm"] * p["nzm"] * p["nrm"] * (p["ntst"] + 1)


class Server:
    """Continuous-batching request loop: one resident Batch per grid shape,
    each advanced by its own thread; admission is global."""

    def __init__(self, tc, hbm_budget=200e9, chunk=25, graph=True, tuning=None, workers=4,
                 max_slots=16, max_cells=400e6, order="arrival", checkpoint_dir=None,
                 checkpoint_every=0):
        # graph=True: each class worker replays its batch's whole step as a
        # hipGraph, captured again whenever jobs join or leave.
        # checkpoint_dir / checkpoint_every: each job's whole state is written
        # to checkpoint_dir/job<index>.tcr (WriteFlowField's role) at the first
        # chunk boundary past every multiple of checkpoint_every steps and
        # when it ends, and at its end job<index>.end records its steps and
        # exit code; run(resume=True) continues each job that has a
        # checkpoint from it, bit for bit the run that wrote it, and reports
        # a job that had already ended (no steps left, or ended by an error
        # code) without running it again.
        if checkpoint_every and not checkpoint_dir:
            raise ValueError("checkpoint_every needs a checkpoint_dir")
        self.checkpoint_dir = checkpoint_dir
        self.checkpoint_every = int(checkpoint_every)
        self.tc = tc
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.max_cells = max_cells
        self.chunk = chunk
        self.graph = graph
        # "arrival": arrived jobs are considered in arrival order;
        # "shortest": the least work (grid points x steps) first, arrival
        # order among equals
        if order not in ("arrival", "shortest"):
            raise ValueError(f"unknown admission order {order!r}")
        self.order = order
        self.tuning = tuning or {}
        self.pool = cf.ThreadPoolExecutor(max_workers=workers)
        self.ahead = workers + 2
        self.cv = threading.Condition()

    def _admitted(self, q):
        with self.cv:
            self.rec[q]["admitted"] = time.perf_counter() - self.t0

    def _resumed(self, q, step):
        with self.cv:
            self.rec[q]["resumed_at"] = step

    def checkpoint_path(self, q):
        return os.path.join(self.checkpoint_dir, f"job{q:04d}.tcr")

    def _checkpoint(self, b, iid, q, end=None):
        # written whole before it replaces the previous one; end = (steps,
        # code) when the job has ended
        path = self.checkpoint_path(q)
        b.write_restart(iid, path + ".tmp")
        os.replace(path + ".tmp", path)
        if end is not None:
            ep = self.end_path(q)
            with open(ep + ".tmp", "w") as f:
                json.dump({"steps": int(end[0]), "code": int(end[1])}, f)
            os.replace(ep + ".tmp", ep)

    def end_path(self, q):
        return os.path.join(self.checkpoint_dir, f"job{q:04d}.end")

    def _ended_before(self, q, p):
        """(steps, code) of a job whose checkpoints say it has nothing left
        to run, else None: its last checkpoint is past NTST, or it is the
        one written when the job ended by an exit code other than 555."""
        ck = self.checkpoint_path(q)
        if not os.path.exists(ck):
            return None
        n = self.tc.restart_info(ck)["ntime"]
        end = None
        if os.path.exists(self.end_path(q)):
            with open(self.end_path(q)) as f:
                end = json.load(f)
            if end["steps"] != n:
                end = None  # from an earlier end, before the job was extended
        if n > p["ntst"]:
            return (n, end["code"] if end else 555)
        if end and end["code"] != 555:
            return (n, end["code"])
        return None

    def _finished(self, q, steps, code, fields):
        with self.cv:
            if q is not None:
                r = self.rec[q]
                r.update(done=time.perf_counter() - self.t0, steps=steps, code=code)
                if fields is not None:
                    r["fields"] = fields
                p0 = self.jobs[q][1]
                self.used -= footprint_bytes(
                    p0, self.tuning.get((p0["nthm"], p0["nzm"], p0["nrm"])))
                self.cells -= r["cells"]
                self.nactive -= 1
            self.events += 1
            self.cv.notify_all()

    def run(self, jobs, keep_fields=False, resume=False):
        """jobs: (arrival s, params, seed).  Returns one record per job:
        arrival, admitted, done (s since start), steps, code, cells; with
        keep_fields, the final (qt, qr, qz, pr); with resume, a job whose
        checkpoint exists continues from it (resumed_at: its step)."""
        if resume and not self.checkpoint_dir:
            raise ValueError("resume needs a checkpoint_dir")
        if self.checkpoint_dir:
            os.makedirs(self.checkpoint_dir, exist_ok=True)
        self.jobs, self.keep_fields = jobs, keep_fields
        self.rec = [None] * len(jobs)
        self.used, self.cells, self.nactive = 0.0, 0, 0
        self.events = 0  # jobs finished (or workers failed) so far
        self.t0 = time.perf_counter()
        workers = {}
        waiting = sorted(range(len(jobs)), key=lambda q: jobs[q][0])
        if resume:
            for q in list(waiting):
                ended = self._ended_before(q, jobs[q][1])
                if ended:
                    p0 = jobs[q][1]
                    self.rec[q] = dict(arrival=jobs[q][0], admitted=None, done=0.0,
                                       steps=ended[0], code=ended[1], resumed_at=ended[0],
                                       cells=p0["nthm"] * p0["nzm"] * p0["nrm"],
                                       steps_total=p0["ntst"] + 1)
                    if keep_fields:
                        self.rec[q]["fields"] = None
                    waiting.remove(q)
        futures = {}
        try:
            while True:
                with self.cv:
                    for w in workers.values():
                        if w.error:
                            raise w.error
                    if not waiting and self.nactive == 0:
                        break
                    seen = self.events
                for q in waiting[: self.ahead]:
                    if q not in futures:
                        futures[q] = self.pool.submit(self._prepare, jobs[q][1], jobs[q][2])
                now = time.perf_counter() - self.t0
                # admission: arrived jobs in the admission order, each if a
                # slot, the HBM budget and the resident-cell target allow
                # (always at least one resident job); later jobs may backfill
                # past one that waits
                arrived = [q for q in waiting if jobs[q][0] <= now]
                if self.order == "shortest":
                    arrived.sort(key=lambda q: (_work(jobs[q][1]), jobs[q][0]))
                # one decision over one snapshot of the resident set: a job
                # finishing meanwhile must not let a later job of the same
                # size pass one that was just turned away
                take = []
                with self.cv:
                    for q in arrived:
                        p0 = jobs[q][1]
                        need = footprint_bytes(
                            p0, self.tuning.get((p0["nthm"], p0["nzm"], p0["nrm"])))
                        c = p0["nthm"] * p0["nzm"] * p0["nrm"]
                        if self.nactive >= self.max_slots:
                            break
                        if self.nactive and (self.used + need > self.budget or
                                             self.cells + c > self.max_cells):
                            continue
                        self.used += need
                        self.cells += c
                        self.nactive += 1
                        take.append((q, c))
                for q, c in take:
                    p0 = jobs[q][1]
                    if q not in futures:
                        futures[q] = self.pool.submit(self._prepare, p0, jobs[q][2])
                    waiting.remove(q)
                    p, fields = futures.pop(q).result()
                    sh = (p["nthm"], p["nzm"], p["nrm"])
                    if sh not in workers:
                        workers[sh] = _ClassWorker(self, sh)
                    with self.cv:
                        self.rec[q] = dict(arrival=jobs[q][0], admitted=None, cells=c,
                                           steps_total=p["ntst"] + 1)
                    ck = self.checkpoint_path(q) if resume else None
                    if ck and not os.path.exists(ck):
                        ck = None
                    workers[sh].admit(q, p, fields, self.tuning.get(sh), ck)
                    del fields
                # sleep until a job finishes or the next arrival; a job that
                # finished during this pass wakes the loop at once
                with self.cv:
                    nxt = min((jobs[q][0] for q in waiting), default=None)
                    timeout = 0.5 if nxt is None else max(0.0, nxt - (time.perf_counter() - self.t0))
                    if self.events == seen:
                        self.cv.wait(min(timeout, 0.5))
        finally:
            for w in workers.values():
                w.close()
        return self.rec

    def _prepare(self, p, seed):
        p = dict(p)
        fields = start_fields(self.tc, p, seed)
        return p, fields
