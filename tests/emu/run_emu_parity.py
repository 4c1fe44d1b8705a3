"""Runs a few parity cases of the product kernel sources compiled for the CPU
(tests/emu, threads + barrier model of a workgroup) against the oracle.
Test infrastructure: executed in a subprocess by tests/test_emu_kernels.py."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cudanavierstokes_amd import _lib  # noqa: E402

_lib._lib = _lib.load(sys.argv[1])

import cudanavierstokes_amd as cns  # noqa: E402
from tests import cases  # noqa: E402
from tests.helpers import initial_fields, oracle_from, state_errors  # noqa: E402

CASES = [
    ("rhs", "tgv", dict(mx=12, my=8, mz=8, stencil_size=2, stencil_visc=2), 0),
    ("rhs", "channel", dict(mx=14, my=8, mz=8, stencil_size=3, stencil_visc=2, periodic_x=0, nonuniform_x=1,
                            forcing=1, Lx=2.0, Ma=1.5, Re=3000.0, viscexp=0.7), 0),
    ("step", "tgv", dict(mx=8, my=8, mz=8, stencil_size=2, stencil_visc=2, scheme=0, check_cfl=1, check_bulk=1), 1),
    ("step", "tgv", dict(mx=8, my=8, mz=8, stencil_size=1, stencil_visc=1, scheme=1), 1),
    ("step", "channel", dict(mx=10, my=8, mz=8, stencil_size=2, stencil_visc=2, scheme=2, periodic_x=0,
                             nonuniform_x=1, forcing=1, Lx=2.0, Ma=1.5, Re=3000.0, viscexp=0.7, check_cfl=1), 1),
]

# FP32 storage against the FP64 oracle (relative L2 <= 1e-5): offset EOS,
# FP32 viscosity law and wall ghosts, compensated RK sums
CASES32 = [
    ("step", "tgv", dict(mx=12, my=12, mz=8, stencil_size=2, stencil_visc=2, scheme=2), 2),
    ("step", "channel", dict(mx=14, my=8, mz=8, stencil_size=3, stencil_visc=2, periodic_x=0, nonuniform_x=1,
                             forcing=1, Lx=2.0, Ma=1.5, Re=3000.0, viscexp=0.7, scheme=0, check_cfl=1), 2),
    ("step", "channel", dict(mx=10, my=8, mz=8, stencil_size=2, stencil_visc=2, periodic_x=0, nonuniform_x=0,
                             Lx=2.0, Ma=0.8, Re=500.0, viscexp=0.7, scheme=1, twall_bot=1.1, twall_top=0.9), 2),
]

# boundary layer (FP64): extrapolated x top and z ends, adiabatic x bottom,
# the blowing band k in [90, 130] (mz = 136, at t = 2.3), chunk lengths that
# the host stretches so that no chunk is shorter than S (mz = 25)
BL = dict(periodic_x=0, nonuniform_x=1, boundary_layer=1, perturbed=1, stretch=5.0, Lx=20.0, Ly=7.0, Lz=50.0,
          Re=1500.0, Pr=0.75, Ma=0.35, viscexp=1.5, amp=0.05)
CASES_BL = [
    ("rhs", "random", dict(BL, mx=16, my=8, mz=25, stencil_size=3, stencil_visc=2), 0),
    ("rhs", "random", dict(BL, mx=16, my=8, mz=136, stencil_size=2, stencil_visc=1, Lz=500.0, time=2.3), 0),
    ("step", "random", dict(BL, mx=16, my=8, mz=20, stencil_size=3, stencil_visc=2, scheme=0, check_cfl=1), 1),
    ("step", "random", dict(BL, mx=16, my=8, mz=24, stencil_size=4, stencil_visc=3, viscexp=1.0, scheme=2), 1),
]

bad = []
for precision, cases_, tol in ((0, CASES, 1e-12), (1, CASES32, 1e-5), (0, CASES_BL, 1e-12)):
    for what, kind, kw, ns in cases_:
        kw = dict(kw)
        t0 = kw.pop("time", None)
        amp = kw.pop("amp", None)
        cfg = cns.Config(**kw)
        orc = oracle_from(cfg)
        f = initial_fields(kind, orc) if amp is None else cases.random_smooth_fields(orc.grid(), orc.cfg, amp=amp)
        orc.set_state(*f)
        sol = cns.Solver(cfg.replace(precision=precision))
        sol.set_state(*f)
        if t0 is not None:
            orc.time = t0
            ctl = sol.control()
            ctl["time"] = t0
            sol.set_control(ctl)
        if what == "rhs":
            e = cases.rel_errors(sol.rhs(), orc.rhs())
        else:
            orc.step(ns)
            sol.step(ns)
            e = state_errors(sol.get_state(), orc.get_state())
        sol.close()
        m = max(v[0] for v in e.values())
        print(what, kind, kw.get("scheme", "-"), "fp32" if precision else "fp64", f"{m:.2e}", flush=True)
        if not m <= tol:
            bad.append((what, kind, precision, m))
print("FAILED" if bad else "OK", bad)
sys.exit(1 if bad else 0)
