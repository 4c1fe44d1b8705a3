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

worst = 0.0
for what, kind, kw, ns in CASES:
    cfg = cns.Config(**kw)
    orc = oracle_from(cfg)
    f = initial_fields(kind, orc)
    orc.set_state(*f)
    sol = cns.Solver(cfg)
    sol.set_state(*f)
    if what == "rhs":
        e = cases.rel_errors(sol.rhs(), orc.rhs())
    else:
        orc.step(ns)
        sol.step(ns)
        e = state_errors(sol.get_state(), orc.get_state())
    sol.close()
    m = max(v[0] for v in e.values())
    worst = max(worst, m)
    print(what, kind, kw.get("scheme", "-"), f"{m:.2e}", flush=True)
print("WORST", worst)
sys.exit(0 if worst <= 1e-12 else 1)
