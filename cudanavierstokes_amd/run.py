"""Command-line driver in the shape of the reference's main program
(src/main.cpp:22-75 with solverWrapper, src/cuda_main.cu:267-325): build the
case, start from the initial condition or restart from field files, then
advance nfiles x nsteps steps, writing the field files after every nsteps.

Output directory layout (the reference's):
  <out>/Grid.txt                i, x, dxi/dx, d2x/dxi2 per x node (src/init.cpp:73-76)
  <out>/prof.txt                wall-normal profiles of calcAvgChan (src/init.cpp:150-208),
                                rewritten at the start and after every file (src/cuda_main.cu:311)
  <out>/fields/{x,y,z}.bin      node coordinates (src/init.cpp:78-88)
  <out>/fields/<c>.%07d.bin     state after each file (c = r, u, v, w, e)
  <out>/fields/ctl.%07d.json    configuration and time-advance control of each file
  <out>/fields/fields.xmf       XDMF index of the field files
  <out>/solution.txt            step, time, par1, par2, dt at every check_bulk-th
                                step (src/cuda_main.cu:306-310; the first column
                                is the global step number)

    python -m cudanavierstokes_amd.run --case tgv --n 64 --order 4 --nfiles 5 --nsteps 100
    python -m cudanavierstokes_amd.run --case channel --restart 5 --nfiles 5
"""
import argparse
import json
import math
import os
import sys
import time

from . import fields as ff
from .solver import Solver, channel, taylor_green


def case_config(a):
    common = dict(order=a.order, scheme=a.scheme, precision=1 if a.precision == "fp32" else 0,
                  check_cfl=a.check_cfl, check_bulk=a.check_bulk, cfl=a.cfl)
    if a.case == "tgv":
        kw = {k: v for k, v in (("Re", a.Re), ("Ma", a.Ma)) if v is not None}
        return taylor_green(n=a.n, **common, **kw)
    kw = {k: v for k, v in (("Re", a.Re), ("Ma", a.Ma)) if v is not None}
    return channel(mx=a.mx, my=a.my, mz=a.mz, **common, **kw)


def write_grid_txt(path, g):
    with open(path, "w") as f:
        for i, (x, xp, xpp) in enumerate(zip(g["x"], g["xp"], g["xpp"])):
            f.write(f"{i} {x:f} {xp:f} {xpp:f}\n")


def write_prof_txt(path, prof):
    cols = [prof[c] for c in ("x", "rm", "um", "vm", "wm", "em", "rf", "uf", "vf", "wf", "ef")]
    with open(path, "w") as f:
        for row in zip(*cols):
            f.write("\t".join(f"{v:f}" for v in row) + "\n")


def print_res(s, g):
    """the resolution report of printRes (src/init.cpp:242-255)"""
    ret = s.friction_reynolds()
    x, y, z = g["x"], g["y"], g["z"]
    mx = len(x)
    print(f"The average friction Reynolds number is: \t {ret:f}")
    print(f"wall-normal wall: \t {(x[0] + x[1]) / 2 * ret:f}")
    print(f"wall-normal center: \t {(x[mx // 2 + 1] - x[mx // 2 - 1]) / 2 * ret:f}")
    print(f"span-wise: \t {(y[1] - y[0]) * ret:f}")
    print(f"stream-wise: \t {(z[1] - z[0]) * ret:f}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description="compressible Navier-Stokes run with field-file output")
    ap.add_argument("--case", choices=("tgv", "channel"), default="tgv")
    ap.add_argument("--n", type=int, default=64, help="Taylor-Green points per side")
    ap.add_argument("--mx", type=int, default=160)
    ap.add_argument("--my", type=int, default=128)
    ap.add_argument("--mz", type=int, default=192)
    ap.add_argument("--order", type=int, default=4, choices=(2, 4, 6, 8))
    ap.add_argument("--scheme", default="rk3", choices=("rk3", "kutta_rk3", "rk4"))
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--Re", type=float, default=None)
    ap.add_argument("--Ma", type=float, default=None)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--check-cfl", type=int, default=10)
    ap.add_argument("--check-bulk", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1, help="channel perturbation seed")
    ap.add_argument("--nfiles", type=int, default=1)
    ap.add_argument("--nsteps", type=int, default=100)
    ap.add_argument("--restart", type=int, default=-1, help="field-file number to restart from (-1: initial condition)")
    ap.add_argument("--out", default=".")
    a = ap.parse_args(argv)

    fdir = os.path.join(a.out, "fields")
    os.makedirs(fdir, exist_ok=True)
    if a.restart >= 0 and os.path.exists(ff.ctl_path(fdir, a.restart)):
        s = ff.solver_from_checkpoint(fdir, a.restart)  # configuration from the checkpoint
    else:
        s = Solver(case_config(a))
        if a.restart >= 0:
            s.read_fields(fdir, a.restart)
        elif a.case == "tgv":
            s.init_taylor_green()
        else:
            s.init_channel(seed=a.seed)
    c = s.config
    g = s.grid()
    write_grid_txt(os.path.join(a.out, "Grid.txt"), g)
    s.write_grid(fdir)
    print_res(s, g)
    write_prof_txt(os.path.join(a.out, "prof.txt"), s.channel_profiles())
    start = max(a.restart, 0)
    if a.restart < 0:
        ff.save_checkpoint(s, fdir, 0)
    with open(os.path.join(a.out, "solution.txt"), "a" if a.restart >= 0 else "w") as sol:
        for file in range(start + 1, start + a.nfiles + 1):
            t0 = time.perf_counter()
            kin, enst, tim = s.step(a.nsteps)
            wall = time.perf_counter() - t0
            ff.save_checkpoint(s, fdir, file)
            write_prof_txt(os.path.join(a.out, "prof.txt"), s.channel_profiles())
            st = s.flow_stats()
            base = (file - 1) * a.nsteps
            for t in range(0, a.nsteps - 1, c.check_bulk):
                dt = tim[t + 1] - tim[t]
                sol.write(f"{base + t + 1} {tim[t]:f} {kin[t]:f} {enst[t]:f} {dt:f}\n")
            sol.flush()
            pts = c.mx * c.my * c.mz * a.nsteps / wall / 1e9
            last = next((i for i in range(a.nsteps - 1, -1, -1) if not math.isnan(kin[i])), a.nsteps - 1)
            print(f"file number: {file}  step: {file * a.nsteps}  time: {tim[-1]:f}  kin: {kin[last]:e}  "
                  f"energy: {enst[last]:e}  kinetic: {st['kinetic']:e}  enstrophy: {st['enstrophy']:e}  "
                  f"({pts:.3f} Gpt-steps/s)", flush=True)
    # index every field file present in the directory (earlier runs included)
    have = sorted({int(f[2:9]) for f in os.listdir(fdir) if f.startswith("r.") and f.endswith(".bin")})
    tmap = {}
    for t in have:
        p = ff.ctl_path(fdir, t)
        if os.path.exists(p):
            with open(p) as f:
                tmap[t] = json.load(f)["control"]["time"]
    ts = [t for t in have if t in tmap]
    ff.write_xdmf(os.path.join(fdir, "fields.xmf"), g["x"], g["y"], g["z"], ts, [tmap[t] for t in ts])
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
