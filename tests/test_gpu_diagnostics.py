"""Device diagnostics against the CPU restatement (oracle/diagnostics.py) on
the state the solver returns: channel profiles of calcAvgChan, the friction
Reynolds number of printRes, kinetic energy / enstrophy / dilatation.
Tolerance: relative 1e-12 (FP64 sums in a different order); FP32 storage
is read back in FP64 on both sides, so the same bound holds."""
import math

import numpy as np
import pytest

import cudanavierstokes_amd as cns
from oracle import diagnostics as D
from oracle import oracle as O
from tests.helpers import flush_print

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-12, atol=0.0):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(a - b)))
    assert err <= rtol * scale + atol, (err, scale)


def _chan(precision=0, **kw):
    return cns.channel(mx=40, my=24, mz=20, order=6, scheme="rk3", Ma=1.2, Re=2000.0, check_cfl=4,
                       precision=precision, **kw)


@pytest.mark.parametrize("precision", [0, 1])
def test_channel_profiles_and_friction_reynolds(precision):
    cfg = _chan(precision)
    flush_print(f"[case] channel diagnostics {cfg}")
    with cns.Solver(cfg) as s:
        s.init_channel(seed=2)
        s.step(5, outputs=False)
        r, u, v, w, e = s.get_state()
        prof = s.channel_profiles()
        again = s.channel_profiles()
        ret = s.friction_reynolds()
        g = s.grid()
    want = D.avg_chan(r, u, v, w, e)
    assert np.array_equal(prof["x"], g["x"])
    for k, a in want.items():
        # fluctuation profiles: compare against the profile's own scale
        _close(prof[k], a, rtol=1e-11 if k.endswith("f") else 1e-12)
        assert np.array_equal(prof[k], again[k]), "repeat calls differ"
    dx = cfg.Lx / cfg.mx  # stretched grid: the computational spacing
    cF = O.coeffs(cfg.stencil_size)[0]
    want_ret = D.friction_reynolds(r, w, cF, dx, g["xp"][0], cfg.Re, cfg.viscexp)
    assert math.isclose(ret, want_ret, rel_tol=1e-12), (ret, want_ret)
    assert ret > 0


STATS = [
    ("tgv_s2", cns.taylor_green(n=24, order=4), "tgv"),
    ("tgv_s4_fp32", cns.taylor_green(n=32, order=8, precision=1), "tgv"),
    ("channel_wall_s3", _chan(0), "channel"),
]


@pytest.mark.parametrize("name,cfg,kind", STATS, ids=[c[0] for c in STATS])
def test_flow_stats(name, cfg, kind):
    flush_print(f"[case] flow stats {name}")
    with cns.Solver(cfg) as s:
        if kind == "tgv":
            s.init_taylor_green()
        else:
            s.init_channel(seed=4)
        s.step(3, outputs=False)
        r, u, v, w, _ = s.get_state()
        st = s.flow_stats()
        g = s.grid()
    h = dict(x=g["x"], xp=g["xp"], dxv=g["dxv"], d_dx=1.0 / (cfg.Lx / cfg.mx if cfg.nonuniform_x else g["x"][1] - g["x"][0]),
             d_dy=1.0 / (g["y"][1] - g["y"][0]), d_dz=1.0 / (g["z"][1] - g["z"][0]))
    cVF = O.coeffs(cfg.stencil_visc)[0]
    want = D.flow_stats(r, u, v, w, h, cVF, cfg.Lx * cfg.Ly * cfg.Lz, wallx=not cfg.periodic_x,
                        nonuni=bool(cfg.nonuniform_x))
    for k in ("kinetic", "enstrophy", "dil2", "rho"):
        assert math.isclose(st[k], want[k], rel_tol=1e-11, abs_tol=1e-300), (k, st[k], want[k])
    if kind == "tgv":
        # the vortex keeps its kinetic energy close to 1/8 over a few steps
        assert abs(st["kinetic"] - 0.125) < 1e-3


@pytest.mark.parametrize("precision", [0, 1])
def test_step_stats_folded_into_the_step_loop(precision):
    """cns_step_stats(6, 2) runs the kernels of three cns_step(2) calls, each
    followed by cns_flow_stats, with the sums on the device: same bits"""
    cfg = cns.taylor_green(n=24, order=4, scheme="rk4", check_cfl=2, precision=precision)
    flush_print(f"[case] step stats {cfg}")
    with cns.Solver(cfg) as a:
        a.init_taylor_green()
        got = a.step_stats(6, 2)
        end_a = a.get_state()
        with pytest.raises(cns.CnsError):
            a.step_stats(5, 2)
    with cns.Solver(cfg) as b:
        b.init_taylor_green()
        want = []
        for _ in range(3):
            b.step(2, outputs=False)
            st = b.flow_stats()
            want.append([st["kinetic"], st["enstrophy"], st["dil2"], st["rho"]])
        end_b = b.get_state()
    assert got.shape == (3, 4)
    assert np.array_equal(got, np.array(want))
    for x, y in zip(end_a, end_b):
        assert np.array_equal(x, y)
    # kinetic energy decays, enstrophy grows at the start of the vortex
    assert got[2, 0] < got[0, 0] and got[2, 1] > got[0, 1]
