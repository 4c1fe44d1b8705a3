#!/usr/bin/env python3
"""Headline benchmark: Taylor-Green vortex, 256^3, 8th-order stencils,
classic RK4, CFL 0.5, FP64, on one MI355X.

A step is one RK4 time step (four fused RHS + stage-update passes) over the
256^3 state, which stays resident in HBM for the whole timed region.  Prints
one JSON line.  The HIP library is the only compute path; the CPU oracle is
timed separately as the cpu_baseline on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpt-RKsteps/s + %HBM roofline, FP64 vs FP32-store; ensemble p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM words per grid point of one classic-RK4 step (4 stages:
# 15 + 25 + 25 + 20 words; state resident, read and write once per stage)
RK4_WORDS = 85
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_rhs_tgv256_o8_rk4_fp64.txt")


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=256, help="grid points per direction")
    ap.add_argument("--order", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: host cores, at most 16)")
    return ap.parse_args()


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def read_traffic():
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-kernel
    calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline(n, order, threads):
    """The FP64 oracle (a CPU restatement of the reference algorithm), one
    RK4 step of the same 256^3 8th-order Taylor-Green case."""
    from oracle import oracle as O
    from tests.cases import tgv_fields
    O.set_threads(threads)
    orc = O.Oracle(mx=n, my=n, mz=n, s=order // 2, sv=order // 2, scheme="rk4", cfl=0.5, check_cfl=10,
                   check_bulk=10, Re=1600.0, Pr=0.71, Ma=0.1)
    orc.set_state(*tgv_fields(orc.grid(), orc.cfg))
    t0 = time.perf_counter()
    orc.step(1)
    dt = time.perf_counter() - t0
    orc.close()
    return {"value": n ** 3 / dt / 1e9, "unit": "Gpt-RKsteps/s", "cores": threads, "kind": "port",
            "sample": f"1 RK4 step of the {n}^3 order-{order} Taylor-Green case on {threads} host threads "
                      f"({dt:.2f} s)"}


def main():
    a = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported (each GPU box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import cudanavierstokes_amd as cns

    n, order = a.n, a.order
    cfg = cns.taylor_green(n=n, order=order, scheme="rk4", cfl=0.5, check_cfl=10, check_bulk=10, Re=1600.0,
                           Pr=0.71, Ma=0.1)
    sol = cns.Solver(cfg)
    sol.set_state(*sol.taylor_green_fields())
    if a.warmup:
        sol.step(a.warmup, outputs=False)
    sol.sync()
    # timed region: HIP events on the solver stream around exactly `steps` steps
    t0 = time.perf_counter()
    ms = sol.time_steps(a.steps)
    wall = time.perf_counter() - t0
    # per-launch durations of the fused RHS + stage-update kernel
    sol.profile_enable(True)
    sol.time_steps(a.steps)
    kms, launches = sol.profile_read()
    sol.profile_enable(False)
    pts = n ** 3
    value = pts * a.steps / (ms / 1e3) / 1e9
    per_launch_ms = kms / max(launches, 1)
    bytes_per_launch = RK4_WORDS * 8 * pts / 4.0
    achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
    traffic = read_traffic()
    step_gbs = RK4_WORDS * 8 * pts * a.steps / (ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gpt-RKsteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex {n}^3, order {order}, classic RK4, CFL 0.5, FP64 storage",
                   "grid": [n, n, n], "order": order, "scheme": "rk4", "cfl": 0.5, "precision": "fp64",
                   "Re": 1600, "Ma": 0.1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_rhs (fused RHS + RK4 stage update)",
                     "kernel_ms_per_launch": round(per_launch_ms, 4), "launches": launches},
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "wall_s": round(wall, 3),
    }
    sol.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, order, a.cpu_threads or host_cores())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
