"""Benchmark: FP64 RK3 steps of the Taylor-Couette solver on one MI355X.

Workload (default): 512 x 512 x 385 (theta x z x r), eta ~ 0.714, ALX3D = 2 pi,
LAMB = 20, 1/NU = 1e4, inner cylinder 1 + 0.1 sin(2 pi 0.05 t), RK3 with
variable dt at CFLMAX = 1.2, wall torque, divergence and velocity maxima
evaluated every step.  The starting field is the laminar Couette profile plus
a seeded 1e-2 perturbation, projected to be divergence free.  A step is one
iteration of the reference main loop: CFL reduction, dt rule, three RK3
substeps, diagnostics.

Prints one JSON line with the throughput in Mpt*RK3-steps/s (grid points
nthm*nzm*nrm times steps per second), the roofline of the dominant kernel
measured with HIP events inside the timed region, and a CPU baseline of the
oracle on this host's cores over a bounded sample of the same workload.
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpt\u00b7RK3-steps/s FP64 at 512\u00b2\u00d7385 + %HBM roofline; sweep jobs/h, p99"
PROFILES = os.path.join(ROOT, "profiles")

# rocprofv3 commands whose output --summarize-profile turns into the committed
# summaries (run from /tmp with TMPDIR=/tmp; OUT is the output folder)
_LEG = "--no-cpu --sweep-steps 0 --served-jobs 0"
# the kernel the roofline object reports (the largest share of the step)
DOMINANT = "helm_fwd_corr"

# counters of the two shader-sequencer passes (per-wave instruction mix;
# where the waves' cycles go)
SQ_PASSES = {
    "SQ_MIX": ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
               "SQ_INSTS_VMEM_WR", "SQ_INSTS_SMEM", "SQ_WAVE_CYCLES"],
    "SQ_STALL": ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY",
                 "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_VMEM", "SQ_LDS_BANK_CONFLICT"],
    # (SQ_ACTIVE_INST_VMEM reads zero for every kernel here and is not summarised)
}


def profile_commands(workload="production_512"):
    """rocprofv3 commands whose output --summarize-profile turns into the
    committed summaries (run from /tmp with TMPDIR=/tmp; OUT is the output
    folder).  Other workloads than the default add --workload; the
    1024x1024x769 one times 3 steps (it reaches the reference's dt floor
    after 12, see configs.large_1024)."""
    w = "" if workload == "production_512" else f" --workload {workload}"
    nst = 3 if workload == "large_1024" else 5
    pmc = ("rocprofv3 --pmc {c} --output-format csv --kernel-include-regex "
           "\"k_helm|k_copy|k_correct|k_fft|k_poisson|k_div|k_diagnostics\" "
           f"-d OUT/{{c}} -o pmc -- python3 bench.py{w} --steps 1 --warmup 1 {_LEG}")
    sq = ("rocprofv3 --pmc {c} --output-format csv --kernel-include-regex "
          "\"k_helm|k_copy|k_correct|k_fft|k_poisson|k_div\" "
          f"-d OUT/{{n}} -o pmc -- python3 bench.py{w} --steps 1 --warmup 1 {_LEG}")
    return {
        "trace": "rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o bench "
                 f"-- python3 bench.py{w} --steps {nst} --warmup 1 {_LEG}",
        "trace_steps": nst,
        "FETCH_SIZE": pmc.format(c="FETCH_SIZE"),
        "WRITE_SIZE": pmc.format(c="WRITE_SIZE"),
        # shader-sequencer counters, two passes of at most eight
        "SQ_MIX": sq.format(n="SQ_MIX", c=" ".join(SQ_PASSES["SQ_MIX"])),
        "SQ_STALL": sq.format(n="SQ_STALL", c=" ".join(SQ_PASSES["SQ_STALL"])),
        # the batched sweep: 2 untimed + 6 timed graph-replayed steps of the 48-instance mix
        "sweep": "rocprofv3 --kernel-trace --stats --output-format csv -d OUT/sweep -o sweep "
                 "-- python3 bench.py --steps 1 --warmup 1 --no-cpu --sweep-steps 6 --served-jobs 0",
    }


PROFILE_COMMANDS = profile_commands()


def kernel_bytes(p):
    """Algorithmic HBM bytes per launch of each kernel: the distinct arrays it
    reads plus those it writes, each counted once (8 B per real point,
    16 B per complex spectral point)."""
    nthm, nzm, nrm = p["nthm"], p["nzm"], p["nrm"]
    N = nthm * nzm * nrm
    S = (nthm // 2 + 1) * nzm * nrm  # complex spectral points
    R, C = 8 * N, 16 * S
    return {
        "explicit_all": 6 * R,       # qt qr qz -> dq dph qcap
        "explicit_qt": 4 * R, "explicit_qr": 4 * R, "explicit_qz": 4 * R,
        "implicit_qt": 8 * R,        # qt pr dq ru1 y | ru1 y qt
        "implicit_qr": 8 * R,        # qr pr dph ru2 y | ru2 y qr
        "implicit_qz": 8 * R,        # qz pr qcap ru3 y | ru3 y qz
        "divergence": 4 * R,         # qt qr qz -> dph
        "fft_theta_fwd": R + C,      # dph -> spec
        "div_fft_theta_fwd": 3 * R + C,  # qt qr qz -> spec (divergence fused)
        "fft_z_fwd": 2 * C, "fft_z_inv": 2 * C,
        "poisson_tridiag": 2 * C,    # spec in place
        "fft_theta_inv": C + R,      # spec -> dph
        "correct_both": 9 * R,       # dph qt qr qz pr | qt qr qz pr
        "correct_velocity": 7 * R, "correct_pressure": 3 * R,
        "max_cfl": 3 * R, "max_div": 3 * R,
        # RK3 substep 1: ro = 0, the history ru is written but not read
        "helm_fwd": 10 * R,          # qt qr qz pr | ru1 ru2 ru3 yt yr yz
        # substeps 2 and 3: the previous substep's correction applied on the fly
        "helm_fwd_corr": 15 * R,     # qt qr qz pr dph ru1 ru2 ru3 | ru1 ru2 ru3 yt yr yz pr
        "helm_bwd": 6 * R,           # yt yr yz (U q + y) | qt qr qz
        "diagnostics": 3 * R,        # qt qr qz -> 10 scalars
        # last correction of a logging step with the diagnostics on the
        # corrected fields: qt qr qz pr dph | qt qr qz pr
        "correct_diag": 9 * R,
    }


def step_bytes_reference(p, n_red):
    """Algorithmic bytes of one RK3 step at the reference's routine
    granularity: 8 N (165 + 24 s + 3 n_red), s = 2 (nzm/2+1)/nzm."""
    N = p["nthm"] * p["nzm"] * p["nrm"]
    s = 2 * (p["nzm"] // 2 + 1) / p["nzm"]
    return 8 * N * (165 + 24 * s + 3 * n_red)


def gpu_leg(tc, p, seed, amp, steps, warmup):
    dev = tc.Solver(p)
    t0 = time.time()
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    dev.set("qt", qt)
    dev.set("qr", qr)
    dev.set("qz", qz)
    del qt, qr, qz
    dev.project()
    print(f"[bench] setup {time.time() - t0:.1f}s, device bytes {dev.device_bytes() / 1e9:.2f} GB",
          file=sys.stderr, flush=True)
    ntime = 0
    for _ in range(warmup):
        code, d = dev.step(ntime)
        ntime += 1
        if code not in (0, 555):
            raise RuntimeError(f"warm-up step {ntime} ended with error code {code}: {d}")
    dev.synchronize()
    # timed region: HIP events around the dominant kernel only, so the wall
    # clock carries no per-kernel event traffic
    dev.kernel_timing(True, only=DOMINANT)
    dev.kernel_time("", reset=True)
    dev.synchronize()
    t0 = time.perf_counter()
    # the K main-loop iterations run in the library (tc_run_steps), each with
    # its CFL, dt rule, RK3 step, logging pass and error codes
    code, n_last, last = dev.run(ntime, steps)
    dev.synchronize()
    t1 = time.perf_counter()
    if code != 0 or n_last != ntime + steps - 1:
        raise RuntimeError(f"timed steps ended at ntime {n_last} with error code {code}: {last}")
    ntime += steps
    dom_ms, dom_cnt = dev.kernel_time(DOMINANT, reset=True)
    # per-kernel table from a few more steps with every kernel timed
    dev.kernel_timing(True)
    dev.kernel_time("", reset=True)
    nt = min(steps, 5)
    for _ in range(nt):
        dev.step(ntime)
        ntime += 1
    dev.synchronize()
    names = list(kernel_bytes(p).keys())
    times = {}
    for n in set(names):
        ms, cnt = dev.kernel_time(n)
        if cnt:
            times[n] = (ms * steps / nt, cnt * steps / nt)  # scaled to the timed steps
    times[DOMINANT] = (dom_ms, dom_cnt)
    # streaming copy of one field, after the timed region: the achievable
    # bandwidth on this box and the reference for hardware byte counters
    dev.kernel_time("", reset=True)
    for _ in range(5):
        dev.copy_field("rhs", "dph")
    ms, cnt = dev.kernel_time("copy", reset=True)
    copy_gbs = 2 * 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"] / (ms / cnt / 1e3) / 1e9
    # STREAM triad (two reads, one write) over three whole fields
    for _ in range(5):
        dev.stream_triad("rhs", "dph", "dq")
    ms, cnt = dev.kernel_time("triad", reset=True)
    triad_gbs = 3 * 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"] / (ms / cnt / 1e3) / 1e9
    dev.kernel_timing(False)
    dev.close()
    return (t1 - t0), times, last, copy_gbs, triad_gbs


def physical_cores():
    """Physical cores (distinct package/core pairs) among the CPUs this
    process may run on, and the number of those logical CPUs."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        top = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(top + "physical_package_id").read().strip(),
                       open(top + "core_id").read().strip()))
        except OSError:
            cores.add(("cpu", c))
    return len(cores), len(cpus)


def cpu_threads():
    """OpenMP threads of the CPU baseline: one per physical core this
    process may use, capped by OMP_NUM_THREADS where the host sets it."""
    phys, _ = physical_cores()
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(phys, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else phys


def cpu_leg(p, seed, amp, steps, warmup):
    """The CPU baseline, run in a child process so that its OpenMP runtime
    starts with one thread per physical core, bound close (OMP_PROC_BIND):
    the oracle rebuilt -O3 -march=native for this host, one projected start,
    `warmup` untimed RK3 steps of the same workload, then `steps` timed
    ones.  Returns the child's JSON record."""
    import subprocess
    import tempfile

    threads = cpu_threads()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as tmp:
        lib = os.path.join(tmp, "liboracle_native.so")
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native",
                        f"NATIVE_OUT={lib}"], check=True)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg-child", lib,
                            "--workload", p["_workload"], "--steps", str(steps),
                            "--warmup", str(warmup)],
                           env=env, check=True, stdout=subprocess.PIPE, text=True)
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    rec["threads"] = threads
    rec["host"] = host_info(rec.pop("triad_gbs"))
    return rec


def cpu_leg_child(lib, workload, steps, warmup):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.use_library(lib)
    tc = importlib.import_module("taylor-couette-afid_amd")
    p = tc.configs.WORKLOADS[workload]()
    seed, amp = tc.configs.SEEDS[workload], tc.configs.AMPLITUDE[workload]
    orc = O.Oracle(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    orc.set("qt", qt)
    orc.set("qr", qr)
    orc.set("qz", qz)
    del qt, qr, qz
    orc.call("project")
    times = []
    for n in range(warmup + steps):
        t0 = time.perf_counter()
        code, _ = orc.step(n)
        if code:
            raise RuntimeError(f"oracle step {n} ended with code {code}")
        if n >= warmup:
            times.append(time.perf_counter() - t0)
        print(f"[bench] cpu baseline step {n}: {time.perf_counter() - t0:.2f} s",
              file=sys.stderr, flush=True)
    orc.close()
    print(json.dumps({"steps": steps, "warmup": warmup, "mean_s": float(np.mean(times)),
                      "p50_s": float(np.median(times)),
                      "triad_gbs": O.stream_triad_gbs(1 << 26, 5)}))


def host_info(triad_gbs):
    """CPU model, sockets and logical CPUs from /proc/cpuinfo and the
    physical cores this process may use, beside the host's STREAM triad
    bandwidth measured on the baseline's OpenMP threads."""
    model, sockets, ncpu = "unknown", set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name":
                    model = v.strip()
                    ncpu += 1
                elif k == "physical id":
                    sockets.add(v.strip())
    except OSError:
        pass
    phys, usable = physical_cores()
    return {"model": model, "sockets": max(len(sockets), 1), "logical_cpus": ncpu,
            "usable_cpus": usable, "usable_physical_cores": phys,
            "triad_gbs": round(triad_gbs, 1)}


def _alone_step_ms(tc, p, fields, steps=8, warmup=2, tuning=None):
    """Per-step time of one job alone on the plain single-instance path
    (tc_step: CFL, dt rule, fused RK3 step, logging diagnostics)."""
    from importlib import import_module
    sw = import_module("taylor-couette-afid_amd.sweep")
    t0 = time.perf_counter()
    s = tc.Solver(p, tuning=tuning)
    sw.load_and_project(s, fields)
    s.synchronize()
    setup = time.perf_counter() - t0
    for n in range(warmup):
        s.step(n)
    s.synchronize()
    t0 = time.perf_counter()
    for n in range(warmup, warmup + steps):
        s.step(n)
    s.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    s.close()
    return ms, setup


def sweep_leg(tc, steps_timed, plain_mpts):
    """Batched sweep of 48 uneven instances: per-step time of the grouped,
    graph-replayed step of all instances, against the same jobs run one at a
    time on the plain path (per-shape step times, whole-job time summed).
    Both sides are reported with their setup (context creation, upload,
    projection) and without it, and beside the bound that the plain
    512x512x385 rate (plain_mpts) puts on any batching of this job mix."""
    from importlib import import_module
    sw = import_module("taylor-couette-afid_amd.sweep")
    jobs = sw.sweep_jobs()
    prepared = []
    for p, seed in jobs:
        p = dict(p)
        prepared.append((p, sw.start_fields(tc, p, seed)))
    key = lambda p: (p["nthm"], p["nzm"], p["nrm"])
    shapes = sorted({key(p) for p, _ in prepared})
    # each side gets its own best schedule: the one-at-a-time jobs the
    # lone-instance autotuner's, the batch the one timed per shape group
    # (tuning time is left out of both sides)
    solo = {sh: tc.autotune(next(q for q, _ in prepared if key(q) == sh), steps=3)
            for sh in shapes}
    tuning = sw.batch_autotune(tc, {sh: [q for q, _ in prepared if key(q) == sh] for sh in shapes},
                               steps=3, base=solo)
    print(f"[bench] sweep: {len(jobs)} instances, shapes {shapes}, tuning alone {solo}, "
          f"batched {tuning}", file=sys.stderr, flush=True)
    # one job at a time on the plain path, per shape
    alone = {}
    for sh in shapes:
        p, f = next((q, f) for q, f in prepared if key(q) == sh)
        alone[sh] = _alone_step_ms(tc, p, f, tuning=solo[sh])
    t0 = time.perf_counter()
    b = tc.Batch()
    for p, f in prepared:
        iid = b.add(p, tuning[key(p)])
        sw.load_and_project(b.solver(iid), f)
    b.groups()
    for i in range(len(jobs)):
        b.solver(i).synchronize()
    setup_batch = time.perf_counter() - t0
    del prepared
    ngroups, launches = b.groups()
    b.advance(2, graph=True, sync_every=2)
    t0 = time.perf_counter()
    b.advance(steps_timed, graph=True, sync_every=steps_timed)
    el = time.perf_counter() - t0
    codes = [b.status(i)[1] for i in range(len(jobs))]
    b.close()
    if any(codes):
        raise RuntimeError(f"sweep instances ended early: {codes}")
    cells = sum(p["nthm"] * p["nzm"] * p["nrm"] for p, _ in jobs)
    steps_job = jobs[0][0]["ntst"] + 1
    ms = 1e3 * el / steps_timed
    alone_steps_s = sum(alone[key(p)][0] * steps_job / 1e3 for p, _ in jobs)
    alone_setup_s = sum(alone[key(p)][1] for p, _ in jobs)
    batch_steps_s = steps_job * ms / 1e3
    with_setup = (alone_steps_s + alone_setup_s) / (batch_steps_s + setup_batch)
    steps_only = alone_steps_s / batch_steps_s
    bound_s = cells * steps_job / (plain_mpts * 1e6)
    return {
        "workload": "48 uneven FP64 instances (64^2x65, 128^2x97, 192x128x129, 256^2x193), "
                    "eta, 1/NU, VIROTEXT/VIROTINT, LAMB seeded; fixed dt at CFL 0.8; 300-step jobs",
        "instances": len(jobs), "groups": ngroups, "launches_per_step": launches,
        "steps_timed": steps_timed, "ms_per_step": round(ms, 3),
        "mpt_steps_per_s": round(cells * steps_timed / el / 1e6, 1),
        "jobs_per_h": round(len(jobs) / (batch_steps_s + setup_batch) * 3600, 1),
        "one_at_a_time_jobs_per_h": round(len(jobs) / (alone_steps_s + alone_setup_s) * 3600, 1),
        "speedup_vs_one_at_a_time": round(with_setup, 2),
        "speedup_steps_only": round(steps_only, 2),
        "setup_s": {"batched": round(setup_batch, 2), "one_at_a_time": round(alone_setup_s, 2)},
        "speedup_bound_at_plain_512_rate": round(alone_steps_s / bound_s, 2),
        "one_at_a_time_ms_per_step": {f"{a}x{b}x{c}": round(v[0], 3) for (a, b, c), v in alone.items()},
        "tuning": {f"{a}x{b}x{c}": {"alone": {k: solo[(a, b, c)][k] for k in ("helm_lpt", "defer")},
                                     "batched": {k: tuning[(a, b, c)][k] for k in ("helm_lpt", "defer")}}
                   for (a, b, c) in shapes},
    }


def served_leg(tc, njobs):
    """Served sweep: Poisson arrivals (2 jobs/s) of 1000-step jobs through the
    continuous-batching server, against the same arrival trace served one
    job at a time on the plain path (run, not modelled)."""
    from importlib import import_module
    sw = import_module("taylor-couette-afid_amd.sweep")
    jobs = sw.served_jobs(njobs)
    shapes = sorted({(p["nthm"], p["nzm"], p["nrm"]) for _, p, _ in jobs})
    tuning = {}
    for sh in shapes:
        _, p, seed = next(j for j in jobs if (j[1]["nthm"], j[1]["nzm"], j[1]["nrm"]) == sh)
        p = dict(p)
        sw.start_fields(tc, p, seed)
        tuning[sh] = tc.autotune(p, steps=3)

    def summary(rec):
        codes = [r["code"] for r in rec]
        if any(c != 555 for c in codes):
            raise RuntimeError(f"served jobs ended with {codes}")
        ttr = np.array([r["done"] - r["arrival"] for r in rec])
        makespan = max(r["done"] for r in rec)
        cell_steps = sum(r["cells"] * r["steps"] for r in rec)
        return {"p50_s": round(float(np.percentile(ttr, 50)), 2),
                "p99_s": round(float(np.percentile(ttr, 99)), 2),
                "jobs_per_h": round(njobs / makespan * 3600, 1),
                "mpt_steps_per_s": round(cell_steps / makespan / 1e6, 1)}, makespan

    srv = sw.Server(tc, tuning=tuning)
    served, m_served = summary(srv.run(jobs))
    print(f"[bench] served {served}", file=sys.stderr, flush=True)
    alone, m_alone = summary(sw.run_one_at_a_time(tc, jobs, tuning))
    print(f"[bench] one at a time {alone}", file=sys.stderr, flush=True)
    out = {"workload": f"{njobs} jobs, Poisson arrivals 2/s, grids 64^2x65..384^2x289, "
                       "eta/Re/rotation ratio/oscillation seeded, 1000 RK3 steps each",
           "jobs": njobs}
    out.update(served)
    out["one_at_a_time"] = alone
    out["speedup_jobs_per_h"] = round(m_alone / m_served, 2)
    return out


def _kernel_short(name):
    """'void tc::k_helm_fwd<true, 2>(tc::GridDev, ...)' -> 'k_helm_fwd<true, 2>'."""
    name = name.replace("void ", "").split("(")[0]
    return name.split("::", 1)[1] if "::" in name else name


# device kernels whose timer name differs from the kernel name
_KERNEL_TIMER = {"k_correct_both2": "correct_both"}


def _timer_of(kernel):
    """Kernel timer name (the kernel_bytes key) of a device kernel name."""
    base = kernel.split("<")[0]
    if base in _KERNEL_TIMER:
        return _KERNEL_TIMER[base]
    if base == "k_helm_fwd" and kernel.replace(" ", "").endswith(",true>"):
        return "helm_fwd_corr"  # k_helm_fwd<UNIFORM, LPT, CORR = true>
    return base[2:] if base.startswith("k_") else base


def summarize_profile(out_dir, workload, p):
    """Turn the CSV output of PROFILE_COMMANDS into the two text summaries
    under profiles/: per-kernel durations, and HBM bytes per dispatch from
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
    counters read for the streaming copy of one field of known size."""
    import csv
    import collections

    kb = kernel_bytes(p)
    cmds = profile_commands(workload)
    N = p["nthm"] * p["nzm"] * p["nrm"]
    R = 8 * N
    copy_bytes = 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"]

    trace_csv = os.path.join(out_dir, "trace", "bench_kernel_stats.csv")
    if os.path.exists(trace_csv):
        rows = list(csv.DictReader(open(trace_csv)))
        lines = [f"# Kernel durations, {workload} ({p['nthm']}x{p['nzm']}x{p['nrm']}, theta x z x r)", "",
                 "Command (from /tmp, TMPDIR=/tmp, OUT = an output folder, repository root as",
                 "the working directory of bench.py):", "", "    " + cmds["trace"], "",
                 f"{cmds['trace_steps']} timed steps + 1 warm-up step + {min(cmds['trace_steps'], 5)} "
                 "more steps with every kernel timed + 5 copies of one field.  Algorithmic GB/s =",
                 "the kernel's distinct input and output arrays (bench.py kernel_bytes) / average",
                 "duration.", "",
                 "| kernel | calls | average us | min us | max us | share % | algorithmic GB | GB/s | "
                 "frac of 8 TB/s |",
                 "|---|---|---|---|---|---|---|---|---|"]
        for r in rows:
            k = _kernel_short(r["Name"])
            t = _timer_of(k)
            avg = float(r["AverageNs"]) / 1e3
            b = kb.get(t) or kb.get(t + "_fwd")
            if t == "copy":
                b = 2 * copy_bytes
            gb = f"{b / 1e9:.3f}" if b else "-"
            gbs = f"{b / (avg * 1e-6) / 1e9:.0f}" if b else "-"
            frac = f"{b / (avg * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f}" if b else "-"
            lines.append(f"| {k} | {r['Calls']} | {avg:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                         f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} | {gb} | {gbs} | "
                         f"{frac} |")
        stats_path = os.path.join(PROFILES, f"all_kernels_{workload}_stats.md")
        open(stats_path, "w").write("\n".join(lines) + "\n")

        print(stats_path)
    if os.path.exists(os.path.join(out_dir, "FETCH_SIZE", "pmc_counter_collection.csv")):
        def load(counter):
            d = collections.defaultdict(list)
            path = os.path.join(out_dir, counter, "pmc_counter_collection.csv")
            for r in csv.DictReader(open(path)):
                d[_kernel_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
            return {k: sum(v) / len(v) for k, v in d.items()}, {k: len(v) for k, v in d.items()}

        fe, cnt = load("FETCH_SIZE")
        wr, _ = load("WRITE_SIZE")
        kf = copy_bytes / (fe["k_copy"] * 1024)
        kw = copy_bytes / (wr["k_copy"] * 1024)
        lines = [f"# HBM traffic per dispatch, {workload} ({p['nthm']}x{p['nzm']}x{p['nrm']})", "",
                 "Commands (from /tmp, TMPDIR=/tmp, OUT = an output folder), one counter per run:", "",
                 "    " + cmds["FETCH_SIZE"], "    " + cmds["WRITE_SIZE"], "",
                 f"Calibration: the streaming copy of one field moves {copy_bytes} B each way; its",
                 f"FETCH_SIZE reads {fe['k_copy']:.0f} KiB and its WRITE_SIZE {wr['k_copy']:.0f} KiB per "
                 f"dispatch, so", f"read bytes = FETCH_SIZE x 1024 x {kf:.4f} and write bytes = WRITE_SIZE "
                 f"x 1024 x {kw:.4f}.", f"One field pass (8 B x {N} cells) = {R / 1e9:.4f} GB.", "",
                 "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per dispatch | "
                 "field passes | algorithmic passes |", "|---|---|---|---|---|---|---|"]
        for k in sorted(fe, key=lambda k: -fe[k] - wr.get(k, 0)):
            t = _timer_of(k)
            b = kb.get(t) or kb.get(t + "_fwd")
            if t == "copy":
                b = 2 * copy_bytes
            hbm = fe[k] * 1024 * kf + wr.get(k, 0.0) * 1024 * kw
            lines.append(f"| {k} | {cnt[k]} | {fe[k]:.0f} | {wr.get(k, 0.0):.0f} | {hbm:.0f} | "
                         f"{hbm / R:.2f} | {(b / R) if b else float('nan'):.2f} |")
        traffic_path = os.path.join(PROFILES, f"all_kernels_{workload}_hbm_traffic.md")
        open(traffic_path, "w").write("\n".join(lines) + "\n")
        print(traffic_path)
    if os.path.exists(os.path.join(out_dir, "SQ_MIX", "pmc_counter_collection.csv")):
        print(summarize_sq(out_dir, workload, p))
    sweep_csv = os.path.join(out_dir, "sweep", "sweep_kernel_trace.csv")
    if os.path.exists(sweep_csv):
        print(summarize_sweep(sweep_csv))


def summarize_sq(out_dir, workload, p):
    """Per-kernel shader-sequencer counters from the two SQ passes: the
    instructions one wave issues, and the share of the waves' cycles spent
    waiting, issuing VALU and vector-memory work.  Cycle counters are in the
    same units (quad cycles), so their ratios are fractions of wave time."""
    import csv
    import collections

    cmds = profile_commands(workload)

    def load(name):
        d = collections.defaultdict(lambda: collections.defaultdict(list))
        path = os.path.join(out_dir, name, "pmc_counter_collection.csv")
        for r in csv.DictReader(open(path)):
            d[_kernel_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in d.items()}

    mix, stall = load("SQ_MIX"), load("SQ_STALL")
    lines = [f"# Shader-sequencer counters, {workload} ({p['nthm']}x{p['nzm']}x{p['nrm']})", "",
             "Commands (from /tmp, TMPDIR=/tmp, OUT = an output folder), one pass each:", "",
             "    " + cmds["SQ_MIX"], "    " + cmds["SQ_STALL"], "",
             "Per wave: instructions issued by one wave (counter / SQ_WAVES, averaged over",
             "dispatches).  Wait and VALU: SQ_WAIT_ANY and SQ_ACTIVE_INST_VALU as fractions",
             "of SQ_WAVE_CYCLES (per wave; with w waves per SIMD the SIMD's VALU is busy",
             "about w times that).  LDS conflict: cycles of SQ_LDS_BANK_CONFLICT per LDS",
             "instruction.  The routine-level kernels of the set-up projection",
             "(k_correct_velocity, k_divergence) appear as well.", "",
             "| kernel | waves | VALU / wave | SALU / wave | LDS / wave | VMEM rd / wave | "
             "VMEM wr / wave | SMEM / wave | wait | VALU active | "
             "LDS conflict cycles / LDS instr |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for k in sorted(mix, key=lambda k: -mix[k].get("SQ_WAVE_CYCLES", 0.0)):
        m, st = mix[k], stall.get(k, {})
        w = m.get("SQ_WAVES", 0.0)
        if w <= 0:
            continue
        per = {c: m.get(c, 0.0) / w for c in SQ_PASSES["SQ_MIX"]}
        wc = st.get("SQ_WAVE_CYCLES", 0.0)
        fr = {c: (st.get(c, 0.0) / wc if wc > 0 else float("nan"))
              for c in ("SQ_WAIT_ANY", "SQ_ACTIVE_INST_VALU")}
        lds_w = m.get("SQ_INSTS_LDS", 0.0) / w
        ws = st.get("SQ_WAVES", 0.0)
        conf = (st.get("SQ_LDS_BANK_CONFLICT", 0.0) / ws / lds_w) if ws > 0 and lds_w > 0 else 0.0
        lines.append(f"| {k} | {w:.0f} | {per['SQ_INSTS_VALU']:.0f} | {per['SQ_INSTS_SALU']:.0f} | "
                     f"{lds_w:.0f} | {per['SQ_INSTS_VMEM_RD']:.0f} | {per['SQ_INSTS_VMEM_WR']:.0f} | "
                     f"{per['SQ_INSTS_SMEM']:.0f} | {fr['SQ_WAIT_ANY']:.2f} | "
                     f"{fr['SQ_ACTIVE_INST_VALU']:.2f} | {conf:.2f} |")
    path = os.path.join(PROFILES, f"sq_counters_{workload}.md")
    open(path, "w").write("\n".join(lines) + "\n")
    return path


def summarize_sweep(path):
    """Grouped kernels of the batched sweep's own steps: the dispatches after
    the last single-instance kernel (autotuning, the one-at-a-time timings
    and the projections come first), i.e. the 8 graph-replayed steps of the
    48-instance mix, with the four shape groups on their own streams."""
    import csv
    import collections

    rows = list(csv.DictReader(open(path)))
    grouped = lambda r: "_g" in _kernel_short(r["Kernel_Name"]).split("<")[0]
    cut = max(int(r["End_Timestamp"]) for r in rows
              if not grouped(r) and "rocclr" not in r["Kernel_Name"])
    g = [r for r in rows if grouped(r) and int(r["Start_Timestamp"]) > cut]
    t0 = min(int(r["Start_Timestamp"]) for r in g)
    t1 = max(int(r["End_Timestamp"]) for r in g)
    busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in g)
    agg = collections.defaultdict(list)
    for r in g:
        agg[(_kernel_short(r["Kernel_Name"]), int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]))].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    steps = 8
    lines = ["# Grouped kernels of the batched sweep (48 instances, 4 shape groups)", "",
             "Command (from /tmp, TMPDIR=/tmp, OUT = an output folder, repository root as",
             "the working directory of bench.py):", "", "    " + PROFILE_COMMANDS["sweep"], "",
             f"The batch's own {steps} steps (2 untimed + 6 timed): the grouped dispatches after "
             "the last", "single-instance kernel.  The four shape groups run on their own "
             "streams inside one", "hipGraph per step, so their kernels overlap:", "",
             f"- wall span {(t1 - t0) / 1e6:.1f} ms, {(t1 - t0) / 1e6 / steps:.2f} ms per step;",
             f"- kernel time summed over dispatches {busy / 1e6:.1f} ms "
             f"({busy / (t1 - t0):.2f} kernels in flight on average);",
             f"- {len(g) // steps} grouped dispatches per step.", "",
             "| kernel | grid x (threads) | grid y (instances, or planes x instances) | calls | "
             "average us | total ms | share of summed % |",
             "|---|---|---|---|---|---|---|"]
    for (k, gx, gy), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| {k} | {gx} | {gy} | {len(v)} | {sum(v) / len(v):.1f} | {sum(v) / 1e3:.2f} | "
                     f"{100 * sum(v) * 1e3 / busy:.1f} |")
    out = os.path.join(PROFILES, "grouped_kernels_sweep48_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    return out


def traffic_of(workload, timer):
    """HBM bytes per dispatch of a kernel from the committed counter summary,
    or None when there is none for this workload."""
    path = os.path.join(PROFILES, f"all_kernels_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.split("|")[1:-1]]
        if len(cells) == 7 and cells[0].startswith("k_"):
            rows.setdefault(_timer_of(cells[0]), cells[4])
    # exact timer first, then the kernel shared by a forward/inverse pair
    for t in (timer, timer.rsplit("_", 1)[0]):
        if t in rows:
            try:
                return int(float(rows[t]))
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="production_512")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="timed RK3 steps of the CPU baseline (after 2 warm-up steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-leg-child", metavar="LIB", help=argparse.SUPPRESS)
    ap.add_argument("--sweep-steps", type=int, default=30,
                    help="timed steps of the batched 48-instance sweep (0: skip)")
    ap.add_argument("--served-jobs", type=int, default=12,
                    help="jobs of the served sweep (0: skip)")
    ap.add_argument("--summarize-profile", metavar="OUT",
                    help="write the profiles/ summaries from rocprofv3 CSV output in OUT and exit")
    a = ap.parse_args()
    if a.cpu_leg_child:
        cpu_leg_child(a.cpu_leg_child, a.workload, a.steps, a.warmup)
        return
    if a.summarize_profile:
        tc = importlib.import_module("taylor-couette-afid_amd")
        summarize_profile(a.summarize_profile, a.workload, tc.configs.WORKLOADS[a.workload]())
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU sharding is a design, not a run)"}))
        sys.exit(2)
    tc = importlib.import_module("taylor-couette-afid_amd")
    p = tc.configs.WORKLOADS[a.workload]()
    seed, amp = tc.configs.SEEDS[a.workload], tc.configs.AMPLITUDE[a.workload]
    npts = p["nthm"] * p["nzm"] * p["nrm"]

    elapsed, ktimes, last, copy_gbs, triad_gbs = gpu_leg(tc, p, seed, amp, a.steps, a.warmup)
    ms_step = 1e3 * elapsed / a.steps
    value = npts * a.steps / elapsed / 1e6

    kb = kernel_bytes(p)
    total_kernel_ms = sum(ms for ms, _ in ktimes.values())
    dom = max((n for n in ktimes if n in kb), key=lambda n: ktimes[n][0])
    if dom != DOMINANT:
        print(f"[bench] note: {dom} outweighs {DOMINANT} in this run", file=sys.stderr)
    dom = DOMINANT
    dom_ms, dom_cnt = ktimes[dom]
    avg_s = dom_ms / dom_cnt / 1e3
    achieved = kb[dom] / avg_s / 1e9
    for n in sorted(ktimes, key=lambda n: -ktimes[n][0]):
        ms, cnt = ktimes[n]
        gbs = (kb[n] / (ms / cnt / 1e3) / 1e9) if n in kb else float("nan")
        print(f"[bench] {n:18s} {ms / a.steps:8.3f} ms/step  {ms / cnt:8.4f} ms/launch  "
              f"{gbs:8.1f} GB/s  {100 * ms / max(total_kernel_ms, 1e-9):5.1f}%", file=sys.stderr)
    n_red = 2 if p.get("tout", 1e30) < 1e-6 else 1
    step_gb_ref = step_bytes_reference(p, n_red) / 1e9
    # the step's own algorithmic bytes: every kernel launch of the timed
    # steps with the bytes its launch moves (the fused kernels' intermediates
    # never reach HBM)
    step_gb = sum(kb[n] * cnt / a.steps for n, (_, cnt) in ktimes.items() if n in kb) / 1e9
    step_gbs = step_gb / (ms_step / 1e3)
    print(f"[bench] {ms_step:.3f} ms/step; fused-path bytes {step_gb:.1f} GB/step = "
          f"{step_gb / (8 * npts / 1e9):.0f} field passes -> {step_gbs:.0f} GB/s "
          f"({step_gbs / HBM_PEAK_GBS:.3f} of {HBM_PEAK_GBS:.0f}); the reference's routine "
          f"granularity would move {step_gb_ref:.1f} GB/step; last {last}", file=sys.stderr)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpt\u00b7RK3-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed Couette start)",
        "config": {"workload": f"{a.workload}: {p['nthm']}x{p['nzm']}x{p['nrm']} "
                               f"(theta x z x r) RK3, variable dt CFL {p['limitcfl']}, "
                               f"torque+divergence every step",
                   "grid_points": npts},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_of(a.workload, dom),
                     "avg_launch_ms": round(dom_ms / dom_cnt, 4),
                     "bytes_per_launch": kb[dom],
                     "copy_kernel_gbs": round(copy_gbs, 1),
                     "triad_kernel_gbs": round(triad_gbs, 1)},
        "step_roofline": {"bound": "hbm", "bytes_per_step": round(step_gb * 1e9),
                          "field_passes": round(step_gb / (8 * npts / 1e9), 1),
                          "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                          # every kernel of the step: average launch (HIP events
                          # on the context's stream: the dominant kernel over the
                          # timed steps, the others over up to 5 more steps),
                          # algorithmic bytes, fraction of the HBM peak
                          "kernels": {
                              n: {"avg_launch_ms": round(ms / cnt, 4),
                                  "launches_per_step": round(cnt / a.steps, 3),
                                  "bytes_per_launch": kb[n],
                                  "achieved_gbs": round(kb[n] / (ms / cnt / 1e3) / 1e9, 1),
                                  "frac": round(kb[n] / (ms / cnt / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
                              for n, (ms, cnt) in sorted(ktimes.items(), key=lambda t: -t[1][0])
                              if n in kb and cnt > 0}},
    }
    if a.sweep_steps > 0:
        out["sweep"] = sweep_leg(tc, a.sweep_steps, value)
        print(f"[bench] sweep {out['sweep']}", file=sys.stderr, flush=True)
    if a.served_jobs > 0:
        out["served"] = served_leg(tc, a.served_jobs)
        print(f"[bench] served {out['served']}", file=sys.stderr, flush=True)
    if not a.no_cpu:
        p["_workload"] = a.workload
        c = cpu_leg(p, seed, amp, a.cpu_steps, 2)
        out["cpu_baseline"] = {"value": round(npts / c["mean_s"] / 1e6, 4),
                               "unit": "Mpt\u00b7RK3-steps/s",
                               "cores": c["threads"], "kind": "port",
                               "sample": f"{c['steps']} timed RK3 steps after {c['warmup']} warm-up "
                                         f"steps of the same {p['nthm']}x{p['nzm']}x{p['nrm']} "
                                         f"workload from the same start; oracle (C restatement, "
                                         f"OpenMP, -O3 -march=native), one thread per physical "
                                         f"core, OMP_PROC_BIND=close",
                               "mean_s_per_step": round(c["mean_s"], 3),
                               "p50_s_per_step": round(c["p50_s"], 3),
                               "value_p50": round(npts / c["p50_s"] / 1e6, 4),
                               "host": c["host"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
