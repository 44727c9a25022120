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
        "fft_z_fwd": 2 * C, "fft_z_inv": 2 * C,
        "poisson_tridiag": 2 * C + C,  # spec in/out, multipliers out/in (8 B each)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="production_512")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
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
        "metric": "Mpt*RK3-steps/s FP64 (512^2x385) + %HBM roofline",
        "value": round(value, 3),
        "unit": "Mpt*steps/s",
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(dom_ms / dom_cnt, 4),
                     "bytes_per_launch": kb[dom],
                     "copy_kernel_gbs": round(copy_gbs, 1)},
    }
    if not a.no_cpu:
        n, t, threads = cpu_leg(p, seed, amp, a.cpu_seconds)
        out["cpu_baseline"] = {"value": round(npts * n / t / 1e6, 4), "unit": "Mpt*steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{n} RK3 step(s) of the same {p['nthm']}x{p['nzm']}x"
                                         f"{p['nrm']} workload from the same start, oracle "
                                         f"(C restatement, OpenMP), {t:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
