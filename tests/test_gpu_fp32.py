"""FP32-storage path (libcns.so with CNS_FP32) against the FP64 oracle.

The conserved variables are stored in FP32 as offsets from a reference state
(rho - rho0, rho E - E0), stencils and fluxes run in FP32, the stage update
runs in FP64 on the compensated pair hi + lo, and metrics and reductions
stay FP64.

Tolerances (relative, per conserved variable; momenta normalised jointly):
  state after N steps   L2 <= 1e-5, Linf <= 1e-4   (with and without compensation)
  one RHS evaluation    L2 <= 1e-3, Linf <= 1e-3
The RHS bound is loose because at low Mach number the mass and energy RHS of
the Taylor-Green vortex are small differences of O(1) fluxes: their FP32
rounding is O(1e-7) of the fluxes but O(1e-4) of the RHS itself.
"""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests import cases
from tests.helpers import assert_close, flush_print, initial_fields, oracle_from, state_errors

pytestmark = pytest.mark.gpu

ST_L2, ST_LINF = 1e-5, 1e-4
RHS_L2, RHS_LINF = 1e-3, 1e-3

CHANNEL = dict(mx=24, my=16, mz=20, stencil_size=3, stencil_visc=3, periodic_x=0, nonuniform_x=1, forcing=1,
               Lx=2.0, Ly=4 * np.pi / 3, Lz=4 * np.pi, Re=3000.0, Pr=0.72, Ma=1.5, viscexp=0.7, check_cfl=4,
               check_bulk=4)

STEP_CASES = [
    ("tgv24_rk4_s4", "tgv", dict(mx=24, my=24, mz=24, stencil_size=4, stencil_visc=4, scheme=2), 20),
    ("tgv16_smr", "tgv", dict(mx=16, my=16, mz=16, stencil_size=2, stencil_visc=2, scheme=0, check_cfl=3), 12),
    ("tgv16_kutta", "tgv", dict(mx=16, my=16, mz=16, stencil_size=2, stencil_visc=2, scheme=1), 10),
    ("channel_smr", "channel", CHANNEL, 10),
    ("random_box_rk4", "random", dict(mx=40, my=20, mz=12, stencil_size=2, stencil_visc=2, Lx=3.0, Ly=2.0,
                                      Lz=1.5, viscexp=0.7, Ma=0.4, Re=300.0, scheme=2), 8),
]


@pytest.mark.parametrize("comp", [1, 0], ids=["compensated", "plain"])
@pytest.mark.parametrize("name,kind,kw,nsteps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_fp32_steps_vs_fp64_oracle(name, kind, kw, nsteps, comp):
    cfg = cns.Config(**kw)
    flush_print(f"[case] fp32 {name} comp={comp} {cfg}")
    orc = oracle_from(cfg)
    f = initial_fields(kind, orc)
    orc.set_state(*f)
    p1o, _, to = orc.step(nsteps)
    so = orc.get_state()
    sol = cns.Solver(cfg.replace(precision=1, compensated=comp))
    sol.set_state(*f)
    p1g, _, tg = sol.step(nsteps)
    errs = state_errors(sol.get_state(), so)
    flush_print(name, comp, errs)
    assert_close(errs, ST_L2, ST_LINF, f"fp32 state {name}")
    # FP64 reductions and the dt chain on FP32 data
    np.testing.assert_allclose(tg, to, rtol=1e-5)
    m = ~np.isnan(p1o)
    np.testing.assert_allclose(p1g[m], p1o[m], rtol=1e-5)
    sol.close()


RHS_CASES = [
    ("tgv_s4", "tgv", dict(mx=24, my=24, mz=24, stencil_size=4, stencil_visc=4)),
    ("channel_s3", "channel", CHANNEL),
    ("random_box", "random", dict(mx=40, my=20, mz=12, stencil_size=2, stencil_visc=2, Lx=3.0, Ly=2.0, Lz=1.5,
                                  viscexp=0.7, Ma=0.4, Re=300.0)),
]


@pytest.mark.parametrize("name,kind,kw", RHS_CASES, ids=[c[0] for c in RHS_CASES])
def test_fp32_rhs_vs_fp64_oracle(name, kind, kw):
    cfg = cns.Config(**kw)
    orc = oracle_from(cfg)
    f = initial_fields(kind, orc)
    orc.set_state(*f)
    sol = cns.Solver(cfg.replace(precision=1))
    sol.set_state(*f)
    errs = cases.rel_errors(sol.rhs(), orc.rhs())
    flush_print(name, errs)
    assert_close(errs, RHS_L2, RHS_LINF, f"fp32 rhs {name}")
    sol.close()


def test_fp32_state_round_trip():
    """set_state -> get_state keeps FP64 values to the compensated pair's
    precision (and to FP32 offset precision without compensation)."""
    cfg = cns.Config(mx=20, my=12, mz=8, stencil_size=2, stencil_visc=2)
    orc = oracle_from(cfg)
    f = initial_fields("random", orc, seed=7)
    for comp, tol in ((1, 1e-13), (0, 1e-6)):
        sol = cns.Solver(cfg.replace(precision=1, compensated=comp))
        sol.set_state(*f)
        errs = state_errors(sol.get_state(), f)
        assert_close(errs, tol, 10 * tol, f"round trip comp={comp}")
        sol.close()


def test_fp32_footprint_is_smaller():
    cfg = cns.Config(mx=32, my=32, mz=32, stencil_size=4, stencil_visc=4, scheme=2)
    b64 = cns.Solver(cfg)
    b32 = cns.Solver(cfg.replace(precision=1, compensated=0))
    b32c = cns.Solver(cfg.replace(precision=1, compensated=1))
    assert b32.device_bytes() < 0.55 * b64.device_bytes()
    assert b32c.device_bytes() < 0.8 * b64.device_bytes()
    for s in (b64, b32, b32c):
        s.close()
