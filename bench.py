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
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-storage run")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of cns_tune")
    ap.add_argument("--other-steps", type=int, default=20, help="steps per other reference configuration (0: skip)")
    ap.add_argument("--err-steps", type=int, default=20, help="steps of the FP32-vs-FP64 error comparison")
    ap.add_argument("--ensemble-steps", type=int, default=100, help="steps of the 64-instance ensemble (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="requests of the seeded serving stream to run (0: skip; the full stream has 1000)")
    ap.add_argument("--cpu-c2-single", action="store_true",
                    help="also time one 256^3 oracle step on a single thread (about half a minute)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (reference point for the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


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


def cpu_topology():
    """CPU model, the CPUs this process may run on, and the physical cores
    among them (lscpu's fields, read from /proc and /sys)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except Exception:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", c))
    return {"model": model, "cpus_allowed": len(cpus), "physical_cores": len(cores)}


def cpu_baseline(full_c2_single=False):
    """The FP64 oracle (a CPU restatement of the reference algorithm; the
    reference has no CPU build) on the host's cores, as planned for the
    baseline: one warm-up step, then the median of the repetitions of the
    step loop alone (steady clock), on 1 thread and on all physical cores
    (OMP_PROC_BIND=close, OMP_PLACES=cores):
      C1  64^3 order 4 low-storage RK3, the full 100 steps (5 repetitions on
          all cores; one repetition of 10 steps on 1 thread)
      C2  256^3 order 8 classic RK4, per-step rate (3 single steps on all
          cores; on 1 thread only with full_c2_single, about half a minute)
      C3  160x128x192 order 6 channel, per-step rate (3 single steps on all
          cores, 1 on 1 thread)
    value: C2 on all cores (the headline workload), Gpt-RKsteps/s."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    from tests.cases import channel_fields, tgv_fields
    topo = cpu_topology()
    allc = max(1, topo["physical_cores"])

    def make(name):
        if name == "C1":
            o = O.Oracle(mx=64, my=64, mz=64, s=2, sv=2, scheme="smr_rk3", cfl=0.5, check_cfl=10, check_bulk=10,
                         Re=1600.0, Pr=0.71, Ma=0.1)
            o.set_state(*tgv_fields(o.grid(), o.cfg))
        elif name == "C2":
            o = O.Oracle(mx=256, my=256, mz=256, s=4, sv=4, scheme="rk4", cfl=0.5, check_cfl=10, check_bulk=10,
                         Re=1600.0, Pr=0.71, Ma=0.1)
            o.set_state(*tgv_fields(o.grid(), o.cfg))
        else:
            o = O.Oracle(mx=160, my=128, mz=192, s=3, sv=3, scheme="smr_rk3", periodic_x=0, nonuniform_x=1,
                         forcing=1, Lx=2.0, Ly=4 * 3.141592653589793 / 3, Lz=4 * 3.141592653589793, Re=3000.0,
                         Pr=0.72, Ma=1.5, viscexp=0.7, stretch=1.8, check_cfl=10, check_bulk=10)
            o.set_state(*channel_fields(o.grid(), o.cfg, seed=1))
        return o

    pts = {"C1": 64 ** 3, "C2": 256 ** 3, "C3": 160 * 128 * 192}

    def run(name, threads, steps, reps):
        O.set_threads(threads)
        o = make(name)
        o.step(1)  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            o.step(steps)
            times.append(time.perf_counter() - t0)
        o.close()
        t = sorted(times)[len(times) // 2]
        return {"config": name, "threads": threads, "steps": steps, "repetitions": reps, "median_s": round(t, 4),
                "gpt_rksteps_per_s": round(pts[name] * steps / t / 1e9, 6)}

    runs = [run("C2", allc, 1, 3), run("C1", allc, 100, 5), run("C3", allc, 1, 3),
            run("C1", 1, 10, 1), run("C3", 1, 1, 1)]
    if full_c2_single:
        runs.append(run("C2", 1, 1, 1))
    head = runs[0]
    return {"value": head["gpt_rksteps_per_s"], "unit": "Gpt-RKsteps/s", "cores": allc, "kind": "port",
            "sample": f"256^3 order-8 Taylor-Green RK4: median of 3 single steps after one warm-up step on "
                      f"{allc} physical cores (the FP64 oracle, OpenMP)",
            "cpu_model": topo["model"], "cpus_allowed": topo["cpus_allowed"], "physical_cores": allc,
            "runs": runs}


def other_configs(cns, steps):
    """Throughput of the other reference configurations (steps with the
    state resident, launch shape tuned): the 64^3 4th-order Taylor-Green case
    with the low-storage RK3, and the 160x128x192 isothermal-wall channel at
    6th order (x wall-normal, stretched; forcing on), FP64 and FP32 storage,
    and the reference's default boundary layer (240x64x2048, order 6/4,
    sponge and wall-normal blowing, FP64) from its Blasius inflow state."""
    from cudanavierstokes_amd import inflow
    out = []
    cases = [("tgv64_o4_rk3_fp64", cns.taylor_green(n=64, order=4, scheme="rk3"), "tgv"),
             ("channel160x128x192_o6_rk3_fp64", cns.channel(mx=160, my=128, mz=192, order=6, scheme="rk3"), "channel"),
             ("channel160x128x192_o6_rk3_fp32",
              cns.channel(mx=160, my=128, mz=192, order=6, scheme="rk3", precision=1), "channel"),
             ("boundary_layer240x64x2048_o6v4_rk3_fp64", cns.boundary_layer(), "blayer")]
    for name, cfg, kind in cases:
        s = cns.Solver(cfg)
        if kind == "tgv":
            s.init_taylor_green()
        elif kind == "blayer":
            inflow.calculate_sponge(s, set_state=True)
        else:
            s.init_channel(seed=1)
        pencil = s.tune()
        s.step(2, outputs=False)
        ms = s.time_steps(steps)
        s.close()
        n = cfg.mx * cfg.my * cfg.mz
        out.append({"workload": name, "grid": [cfg.mx, cfg.my, cfg.mz], "steps": steps, "pencil": pencil,
                    "ms_per_step": round(ms / steps, 4), "gpt_rksteps_per_s": round(n * steps / ms / 1e6, 4)})
    return out


def copy_calibration(mib, reps=5):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE per
    dispatch, against mib MiB read and written, scales the counters read for
    the solver's kernels."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "dispatches": reps}), flush=True)


def ensemble_measure(cns, steps):
    """64 seeded Taylor-Green instances (32^3..128^3, Re 400-3200, Mach
    0.1-0.5), 6th order, RK4, FP32 storage: each alone in turn with its own
    tuned launch shape (cns_tune), then all in one tuned batch (one launch
    per RK stage, graph replay)."""
    from cudanavierstokes_amd.workloads import ensemble
    inst = ensemble(n_instances=64, seed=4, steps=steps)
    work = sum(i["config"].mx * i["config"].my * i["config"].mz for i in inst) * steps
    seq_ms = 0.0
    for i in inst:
        s = cns.Solver(i["config"])
        s.set_state(*s.taylor_green_fields())
        s.tune()
        s.step(1, outputs=False)
        seq_ms += s.time_steps(steps)
        s.close()
    sols = []
    for i in inst:
        s = cns.Solver(i["config"])
        s.set_state(*s.taylor_green_fields())
        sols.append(s)
    b = cns.Batch(sols)
    t0 = time.perf_counter()
    groups = b.tune()
    tune_s = time.perf_counter() - t0
    b.step(1)
    bat_ms = b.time_steps(steps)
    b.close()
    for s in sols:
        s.close()
    seq = work / (seq_ms / 1e3) / 1e9
    bat = work / (bat_ms / 1e3) / 1e9
    return {"workload": "64 Taylor-Green instances, sizes {32,48,64,96,128}^3 (seed 4), order 6, RK4, "
                        f"FP32 storage, {steps} steps each; sequential: each alone, launch shape tuned, timed with "
                        "cns_time_steps (bulk statistics every 10 steps included, which the batch does not compute)",
            "sequential_gpt_steps_per_s": round(seq, 4), "batched_gpt_steps_per_s": round(bat, 4),
            "speedup": round(bat / seq, 3), "sequential_ms": round(seq_ms, 2), "batched_ms": round(bat_ms, 2),
            "tune_s": round(tune_s, 3), "groups": groups}


def serve_measure(n_requests):
    """The first n_requests of the seeded serving stream (Poisson, 20/s; 70%
    Taylor-Green 48^3-160^3, 30% channel; 6th order, RK3, FP32) on the wall
    clock, continuous batching within a 200 GB budget."""
    from cudanavierstokes_amd.serve import Server
    from cudanavierstokes_amd.workloads import serving
    reqs = serving(n_requests=n_requests, seed=5, rate=20.0)
    _, summary = Server(budget_bytes=200e9, tick_steps=10).run(reqs, realtime=True)
    summary["workload"] = (f"{'all' if n_requests >= 1000 else f'first {n_requests} of'} 1000 seeded requests "
                           "(Poisson 20/s), 70% Taylor-Green 48^3-160^3 / 30% channel, order 6, RK3, FP32 storage, "
                           "200 GB admission budget, on the wall clock")
    return summary


def main():
    a = parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported (each GPU box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import cudanavierstokes_amd as cns

    n, order = a.n, a.order
    pts = n ** 3
    base = cns.taylor_green(n=n, order=order, scheme="rk4", cfl=0.5, check_cfl=10, check_bulk=10, Re=1600.0,
                            Pr=0.71, Ma=0.1)

    def measure(cfg, wsize):
        """Throughput of `steps` RK4 steps and the per-launch time of the fused
        kernel (HIP events on the solver stream)."""
        sol = cns.Solver(cfg)
        sol.set_state(*sol.taylor_green_fields())
        # launch shape: a few timed steps per pencil length, state restored
        pencil = sol.tune() if not a.no_tune else None
        if a.warmup:
            sol.step(a.warmup, outputs=False)
        sol.sync()
        # timed region: HIP events on the solver stream around exactly `steps` steps
        t0 = time.perf_counter()
        ms = sol.time_steps(a.steps)
        wall = time.perf_counter() - t0
        sol.profile_enable(True)
        sol.time_steps(a.steps)
        kms, launches = sol.profile_read()
        sol.profile_enable(False)
        sol.close()
        per_launch_ms = kms / max(launches, 1)
        bytes_per_launch = RK4_WORDS * wsize * pts / 4.0
        return {"ms": ms, "wall": wall, "value": pts * a.steps / (ms / 1e3) / 1e9, "pencil": pencil,
                "per_launch_ms": per_launch_ms, "launches": launches,
                "achieved": bytes_per_launch / (per_launch_ms / 1e3) / 1e9,
                "step_gbs": RK4_WORDS * wsize * pts * a.steps / (ms / 1e3) / 1e9}

    def fp32_error(nsteps):
        """Relative L2 / Linf per conserved variable of the FP32-storage state
        against the FP64 state after the same nsteps from the same fields."""
        states = []
        for cfg in (base, base.replace(precision=1, compensated=1)):
            sol = cns.Solver(cfg)
            sol.set_state(*sol.taylor_green_fields())
            sol.step(nsteps, outputs=False)
            r, u, v, w, e = sol.get_state()
            states.append([r, r * u, r * v, r * w, e])
            sol.close()
        ref, got = states
        mom = sum(float((ref[q].astype("float64") ** 2).sum()) for q in (1, 2, 3)) ** 0.5
        mom_inf = max(float(abs(ref[q]).max()) for q in (1, 2, 3))
        out = {}
        for q, nm in enumerate(["r", "ru", "rv", "rw", "re"]):
            d = got[q] - ref[q]
            n2 = mom if q in (1, 2, 3) else float((ref[q] ** 2).sum()) ** 0.5
            ni = mom_inf if q in (1, 2, 3) else float(abs(ref[q]).max())
            out[nm] = [float((d ** 2).sum()) ** 0.5 / n2, float(abs(d).max()) / ni]
        return out

    m64 = measure(base, 8)
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": round(m64["value"], 4),
        "unit": "Gpt-RKsteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(m64["ms"] / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex {n}^3, order {order}, classic RK4, CFL 0.5, FP64 storage",
                   "grid": [n, n, n], "order": order, "scheme": "rk4", "cfl": 0.5, "precision": "fp64",
                   "Re": 1600, "Ma": 0.1},
        "roofline": {"bound": "hbm", "achieved": round(m64["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(m64["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per k_rhs launch, scaled by a copy "
                                       "kernel, from profiles/traffic_k_rhs_tgv256_o8_rk4_fp64.txt (not this run)",
                     "kernel": "k_rhs (fused RHS + RK4 stage update)",
                     "kernel_ms_per_launch": round(m64["per_launch_ms"], 4), "launches": m64["launches"],
                     "pencil": m64["pencil"]},
        "step_hbm_frac": round(m64["step_gbs"] / HBM_PEAK_GBS, 4),
        "wall_s": round(m64["wall"], 3),
    }
    if not a.no_fp32:
        m32 = measure(base.replace(precision=1, compensated=1), 4)
        out["fp32"] = {
            "workload": "same case, FP32 storage (offsets from the mean state), compensated RK sums, "
                        "FP64 metrics and reductions",
            "value": round(m32["value"], 4), "unit": "Gpt-RKsteps/s",
            "ms_per_step": round(m32["ms"] / a.steps, 4),
            "speedup_vs_fp64": round(m32["value"] / m64["value"], 3),
            "roofline": {"bound": "hbm", "achieved": round(m32["achieved"], 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(m32["achieved"] / HBM_PEAK_GBS, 4),
                         "kernel_ms_per_launch": round(m32["per_launch_ms"], 4), "pencil": m32["pencil"]},
            "rel_err_vs_fp64": {"steps": a.err_steps, "l2_linf": fp32_error(a.err_steps)},
        }
    if a.ensemble_steps > 0:
        out["ensemble"] = ensemble_measure(cns, a.ensemble_steps)
    if a.serve_requests > 0:
        out["serving"] = serve_measure(a.serve_requests)
    if a.other_steps > 0:
        out["other_configs"] = other_configs(cns, a.other_steps)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_c2_single)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
