"""Continuous-batching request loop for many simulations on one GPU.

Requests (a configuration, an initial-condition kind, a step count and an
arrival time) are admitted first-come first-served while the predicted device
footprint of the admitted states stays within a budget; admitted states stay
resident until their last step.  Every tick advances all admitted requests
together with one batch (cns_batch_*: one launch per RK stage per kernel
group, one replayed graph per step), by the smallest remaining step count or
the tick length, whichever is less; finished requests leave (a worker
thread reads their results while the next ticks run, and their solvers are
closed between ticks; the library keeps their device blocks cached for the
next requests of the same sizes), waiting ones are admitted, and the batch is
rebuilt for the new membership.  Each request's
result equals its lone cns_step result to rounding (the batch keeps every
member's step count, so the CFL checks fall on the same steps).

A request may start from a checkpoint instead of its initial condition
(restart=dict(dir=..., timestep=...): field files plus the time-advance
control, see fields.save_checkpoint), and with checkpoint_dir set every
finished request leaves one in <checkpoint_dir>/<id>/, numbered by the steps
taken (plus the restart timestep), so a long run can be served in parts.  The
checkpoint carries the CFL check phase, so a resumed request continues as the
uninterrupted one would, to rounding.

Run as a module for the seeded serving workload:
    python -m cudanavierstokes_amd.serve --requests 1000 --rate 20
"""
import argparse
import collections
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .fields import load_checkpoint, save_checkpoint
from .solver import Batch, Solver, release_cache
from .workloads import init_on_device, points, serving

_SETS = {0: (2, 2, 2), 1: (3, 1, 2), 2: (4, 2, 1)}  # scheme -> (state sets, compensated sets, rhs sets)


def footprint(config):
    """Device bytes a solver of this configuration allocates (the large
    arrays; metrics and reduction scratch add well under a megabyte)."""
    n = points(config)
    sets, lo, rhs = _SETS[int(config.scheme)]
    fp32 = int(config.precision) == 1
    esz = 4 if fp32 else 8
    fields = 5 * sets + (5 * lo if fp32 and config.compensated else 0) + 5 * rhs + 1
    small = 8 * (config.mx * (2 * config.stencil_visc + 4) + config.mz + 5 * config.mx * config.mz + 4 * 1024) + 4096
    return fields * n * esz + small


class Server:
    def __init__(self, budget_bytes=200e9, tick_steps=10, max_members=512, tune=False, keep_results=False,
                 checkpoint_dir=None):
        self.budget = float(budget_bytes)
        self.ckpt = checkpoint_dir
        self.tick = int(tick_steps)
        self.max_members = int(max_members)
        self.tune = tune
        self.keep = keep_results

    def run(self, requests, realtime=True, log=None):
        """Serve `requests` (dicts with id, arrival [s], config, kind, steps,
        seed).  realtime: arrivals follow the wall clock; otherwise, when the
        GPU is idle, the clock jumps to the next arrival.  Returns one record
        per request and the run's summary."""
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = collections.deque()
        active = {}    # admitted and in the current batch
        joining = {}   # admitted while the batch runs; they join at the next tick
        done = []
        used = [0.0]
        batch = None
        shift = [0.0]
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0 + shift[0]  # noqa: E731
        busy = 0.0
        ticks = 0
        cost = collections.Counter()  # host seconds per phase

        def admit():
            now = clock()
            while pending and pending[0]["arrival"] <= now:
                waiting.append(pending.popleft())
            while waiting and len(active) + len(joining) < self.max_members:
                r = waiting[0]
                need = footprint(r["config"])
                if need > self.budget:
                    raise ValueError(f"request {r['id']} needs {need:.3g} B, above the budget")
                if used[0] + need > self.budget:
                    break
                waiting.popleft()
                ta = time.perf_counter()
                s = Solver(r["config"])
                if "restart" in r:
                    load_checkpoint(s, r["restart"]["dir"], r["restart"]["timestep"])
                else:
                    init_on_device(s, r["kind"], r.get("seed", 1))
                cost["admit_s"] += time.perf_counter() - ta
                joining[r["id"]] = dict(req=r, solver=s, left=int(r["steps"]), admitted=clock(), bytes=need)
                used[0] += need

        # Finished requests: their results are read (cns_get_state on their
        # own streams) by a worker thread while the next ticks run; their
        # solvers are closed on this thread between ticks (the library keeps
        # their device blocks for the next solvers of the same sizes).
        reader = ThreadPoolExecutor(max_workers=2)
        reading = []  # (future, entry)

        def read_result(a):
            r = a["req"]
            state = a["solver"].get_state()
            t_done = clock()
            rec = dict(id=r["id"], kind=r["kind"], arrival=r["arrival"], admitted=a["admitted"],
                       done=t_done, tts=t_done - r["arrival"], steps=int(r["steps"]),
                       points=points(r["config"]))
            if self.keep:
                rec["state"] = state
            if self.ckpt is not None:
                ts = int(r["steps"]) + (int(r["restart"]["timestep"]) if "restart" in r else 0)
                d = os.path.join(self.ckpt, str(r["id"]))
                save_checkpoint(a["solver"], d, ts)
                rec["checkpoint"] = dict(dir=d, timestep=ts)
            return rec

        def release(block):
            """Close the solvers whose results are read (all of them with
            block); frees their share of the budget."""
            tf = time.perf_counter()
            keep = []
            for fut, a in reading:
                if block or fut.done():
                    rec = fut.result()
                    done.append(rec)
                    a["solver"].close()
                    used[0] -= a["bytes"]
                    if log:
                        log(rec)
                else:
                    keep.append((fut, a))
            reading[:] = keep
            cost["release_s"] += time.perf_counter() - tf

        finished_now = []
        try:
            while pending or waiting or active or joining or reading or finished_now:
                for a in finished_now:
                    reading.append((reader.submit(read_result, a), a))
                finished_now = []
                admit()
                if joining:
                    active.update(joining)
                    joining.clear()
                    if batch is not None:
                        batch.close()
                        batch = None
                if not active:
                    if reading:
                        release(True)  # frees memory: admit again before waiting
                        continue
                    if pending:
                        wait = pending[0]["arrival"] - clock()
                        if wait > 0:
                            if realtime:
                                time.sleep(wait)
                            else:
                                shift[0] += wait
                    continue
                if batch is None:
                    tb = time.perf_counter()
                    batch = Batch([a["solver"] for a in active.values()], tune=self.tune)
                    cost["batch_build_s"] += time.perf_counter() - tb
                k = min(self.tick, min(a["left"] for a in active.values()))
                tb = time.perf_counter()
                batch.launch(k)
                admit()  # newcomers' state upload overlaps the tick
                batch.wait()
                busy += time.perf_counter() - tb
                ticks += 1
                release(False)
                finished = [rid for rid, a in active.items() if a["left"] == k]
                for a in active.values():
                    a["left"] -= k
                if finished:
                    batch.close()
                    batch = None
                    finished_now = [active.pop(rid) for rid in finished]
        finally:
            reader.shutdown(wait=True)
        if batch is not None:
            batch.close()
        wall = clock()
        release_cache()  # closed solvers' device blocks stay cached for reuse until here
        out = summarize(done, wall, busy, ticks)
        out.update({k: round(v, 3) for k, v in cost.items()})
        return done, out


def summarize(done, wall, busy=None, ticks=None):
    tts = np.array([d["tts"] for d in done]) if done else np.zeros(1)
    work = float(sum(d["points"] * d["steps"] for d in done))
    out = {"requests": len(done), "wall_s": round(wall, 3),
           "p50_s": round(float(np.percentile(tts, 50)), 4), "p99_s": round(float(np.percentile(tts, 99)), 4),
           "mean_s": round(float(tts.mean()), 4), "point_steps": work,
           "aggregate_gpt_steps_per_s": round(work / max(wall, 1e-9) / 1e9, 4)}
    if busy is not None:
        out["tick_s"] = round(busy, 3)
        out["tick_gpt_steps_per_s"] = round(work / max(busy, 1e-9) / 1e9, 4)
        out["ticks"] = ticks
    return out


def main():
    ap = argparse.ArgumentParser(description="seeded serving workload on one GPU")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--tick", type=int, default=10)
    ap.add_argument("--no-realtime", action="store_true")
    a = ap.parse_args()
    reqs = serving(n_requests=a.requests, seed=a.seed, rate=a.rate)
    done, summary = Server(budget_bytes=a.budget_gb * 1e9, tick_steps=a.tick).run(reqs, realtime=not a.no_realtime)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
