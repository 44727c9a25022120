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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpt\u00b7RK3-steps/s FP64 at 512\u00b2\u00d7385 + %HBM roofline; sweep jobs/h, p99"
PROFILES = os.path.join(ROOT, "profiles")

# rocprofv3 commands whose output --summarize-profile turns into the committed
# summaries (run from /tmp with TMPDIR=/tmp; OUT is the output folder)
PROFILE_COMMANDS = {
    "trace": "rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o bench "
             "-- python3 bench.py --steps 5 --warmup 1 --no-cpu",
    "FETCH_SIZE": "rocprofv3 --pmc FETCH_SIZE --output-format csv --kernel-include-regex "
                  "\"k_helm|k_copy|k_correct|k_fft|k_poisson|k_divergence|k_diagnostics\" "
                  "-d OUT/FETCH_SIZE -o pmc -- python3 bench.py --steps 1 --warmup 1 --no-cpu",
    "WRITE_SIZE": "rocprofv3 --pmc WRITE_SIZE --output-format csv --kernel-include-regex "
                  "\"k_helm|k_copy|k_correct|k_fft|k_poisson|k_divergence|k_diagnostics\" "
                  "-d OUT/WRITE_SIZE -o pmc -- python3 bench.py --steps 1 --warmup 1 --no-cpu",
}


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
        "helm_fwd": 13 * R,          # qt qr qz pr ru1 ru2 ru3 | ru1 ru2 ru3 yt yr yz
        "helm_bwd": 9 * R,           # yt yr yz qt qr qz | qt qr qz
        "diagnostics": 3 * R,        # qt qr qz -> 10 scalars
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
    dev.kernel_timing(True)
    dev.kernel_time("", reset=True)
    dev.synchronize()
    t0 = time.perf_counter()
    last = None
    for _ in range(steps):
        code, last = dev.step(ntime)
        ntime += 1
        if code not in (0, 555, 333):
            raise RuntimeError(f"step {ntime} ended with error code {code}: {last}")
    dev.synchronize()
    t1 = time.perf_counter()
    names = list(kernel_bytes(p).keys())
    times = {}
    for n in set(names):
        ms, cnt = dev.kernel_time(n)
        if cnt:
            times[n] = (ms, cnt)
    # streaming copy of one field, after the timed region: the achievable
    # bandwidth on this box and the reference for hardware byte counters
    dev.kernel_time("", reset=True)
    for _ in range(5):
        dev.copy_field("rhs", "dph")
    ms, cnt = dev.kernel_time("copy", reset=True)
    dev.kernel_timing(False)
    copy_gbs = 2 * 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"] / (ms / cnt / 1e3) / 1e9
    dev.close()
    return (t1 - t0), times, last, copy_gbs


def cpu_leg(p, seed, amp, max_seconds):
    """Oracle (CPU restatement) on this host: one projection-free start, then
    RK3 steps of the same workload until max_seconds of stepping."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    tc = importlib.import_module("taylor-couette-afid_amd")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc = O.Oracle(p)
    qt, qr, qz = tc.init_perturbed_couette(p, seed, amp)
    orc.set("qt", qt)
    orc.set("qr", qr)
    orc.set("qz", qz)
    del qt, qr, qz
    orc.call("project")
    n, t_steps = 0, 0.0
    while True:
        t0 = time.perf_counter()
        orc.step(n)
        t_steps += time.perf_counter() - t0
        n += 1
        if t_steps >= max_seconds or n >= 50:
            break
    orc.close()
    return n, t_steps, threads


def _kernel_short(name):
    """'void tc::k_helm_fwd<true, 2>(tc::GridDev, ...)' -> 'k_helm_fwd<true, 2>'."""
    name = name.replace("void ", "").split("(")[0]
    return name.split("::", 1)[1] if "::" in name else name


def _timer_of(kernel):
    """Kernel timer name (the kernel_bytes key) of a device kernel name."""
    base = kernel.split("<")[0]
    return base[2:] if base.startswith("k_") else base


def summarize_profile(out_dir, workload, p):
    """Turn the CSV output of PROFILE_COMMANDS into the two text summaries
    under profiles/: per-kernel durations, and HBM bytes per dispatch from
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
    counters read for the streaming copy of one field of known size."""
    import csv
    import collections

    kb = kernel_bytes(p)
    N = p["nthm"] * p["nzm"] * p["nrm"]
    R = 8 * N
    copy_bytes = 8 * (p["nrm"] + 1) * p["nzm"] * p["nthm"]

    rows = list(csv.DictReader(open(os.path.join(out_dir, "trace", "bench_kernel_stats.csv"))))
    lines = [f"# Kernel durations, {workload} ({p['nthm']}x{p['nzm']}x{p['nrm']}, theta x z x r)", "",
             "Command (from /tmp, TMPDIR=/tmp, OUT = an output folder, repository root as",
             "the working directory of bench.py):", "", "    " + PROFILE_COMMANDS["trace"], "",
             "5 timed steps + 1 warm-up step + 5 copies of one field.  Algorithmic GB/s =",
             "the kernel's distinct input and output arrays (bench.py kernel_bytes) / average",
             "duration.", "",
             "| kernel | calls | average us | min us | max us | share % | algorithmic GB | GB/s |",
             "|---|---|---|---|---|---|---|---|"]
    for r in rows:
        k = _kernel_short(r["Name"])
        t = _timer_of(k)
        avg = float(r["AverageNs"]) / 1e3
        b = kb.get(t) or kb.get(t + "_fwd")
        if t == "copy":
            b = 2 * copy_bytes
        gb = f"{b / 1e9:.3f}" if b else "-"
        gbs = f"{b / (avg * 1e-6) / 1e9:.0f}" if b else "-"
        lines.append(f"| {k} | {r['Calls']} | {avg:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} | {gb} | {gbs} |")
    stats_path = os.path.join(PROFILES, f"all_kernels_{workload}_stats.md")
    open(stats_path, "w").write("\n".join(lines) + "\n")

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
             "    " + PROFILE_COMMANDS["FETCH_SIZE"], "    " + PROFILE_COMMANDS["WRITE_SIZE"], "",
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
    print(stats_path)
    print(traffic_path)


def traffic_of(workload, timer):
    """HBM bytes per dispatch of a kernel from the committed counter summary,
    or None when there is none for this workload."""
    path = os.path.join(PROFILES, f"all_kernels_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.split("|")[1:-1]]
        if len(cells) == 7 and cells[0].startswith("k_") and _timer_of(cells[0]) in (timer, timer.rsplit("_", 1)[0]):
            try:
                return int(float(cells[4]))
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="production_512")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--summarize-profile", metavar="OUT",
                    help="write the profiles/ summaries from rocprofv3 CSV output in OUT and exit")
    a = ap.parse_args()
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

    elapsed, ktimes, last, copy_gbs = gpu_leg(tc, p, seed, amp, a.steps, a.warmup)
    ms_step = 1e3 * elapsed / a.steps
    value = npts * a.steps / elapsed / 1e6

    kb = kernel_bytes(p)
    total_kernel_ms = sum(ms for ms, _ in ktimes.values())
    dom = max((n for n in ktimes if n in kb), key=lambda n: ktimes[n][0])
    dom_ms, dom_cnt = ktimes[dom]
    avg_s = dom_ms / dom_cnt / 1e3
    achieved = kb[dom] / avg_s / 1e9
    for n in sorted(ktimes, key=lambda n: -ktimes[n][0]):
        ms, cnt = ktimes[n]
        gbs = (kb[n] / (ms / cnt / 1e3) / 1e9) if n in kb else float("nan")
        print(f"[bench] {n:18s} {ms / a.steps:8.3f} ms/step  {ms / cnt:8.4f} ms/launch  "
              f"{gbs:8.1f} GB/s  {100 * ms / max(total_kernel_ms, 1e-9):5.1f}%", file=sys.stderr)
    n_red = 2 if p.get("tout", 1e30) < 1e-6 else 1
    step_gb = step_bytes_reference(p, n_red) / 1e9
    print(f"[bench] {ms_step:.3f} ms/step; reference-granularity bytes {step_gb:.1f} GB/step -> "
          f"{step_gb / (ms_step / 1e3):.0f} GB/s equivalent; last {last}", file=sys.stderr)

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
                     "copy_kernel_gbs": round(copy_gbs, 1)},
    }
    if not a.no_cpu:
        n, t, threads = cpu_leg(p, seed, amp, a.cpu_seconds)
        out["cpu_baseline"] = {"value": round(npts * n / t / 1e6, 4), "unit": "Mpt\u00b7RK3-steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{n} RK3 step(s) of the same {p['nthm']}x{p['nzm']}x"
                                         f"{p['nrm']} workload from the same start, oracle "
                                         f"(C restatement, OpenMP), {t:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
