"""The CPU restatement of the state diagnostics (oracle/diagnostics.py)
against the reference's own loop nests and against closed forms."""
import math

import numpy as np

from oracle import diagnostics as D
from oracle import oracle as O


def _rand(shape, seed, lo=0.5, hi=1.5):
    return np.random.default_rng(seed).uniform(lo, hi, shape)


def test_avg_chan_numpy_equals_reference_loops():
    sh = (5, 4, 6)
    r, u, v, w, e = (_rand(sh, s) for s in range(5))
    a, b = D.avg_chan_loops(r, u, v, w, e), D.avg_chan(r, u, v, w, e)
    for k in a:
        assert np.allclose(a[k], b[k], rtol=1e-13, atol=1e-16), k


def test_avg_chan_of_plane_uniform_fields():
    mx = 7
    prof = np.linspace(1.0, 2.0, mx)
    f = [np.broadcast_to(c * prof, (3, 5, mx)).copy() for c in (1.0, 0.3, -0.2, 1.7, 2.5)]
    r, ru, rv, rw, e = f
    out = D.avg_chan(r, ru / r, rv / r, rw / r, e)
    assert np.allclose(out["rm"], prof, rtol=1e-15)
    assert np.allclose(out["um"], 0.3, rtol=1e-14) and np.allclose(out["wm"], 1.7, rtol=1e-14)
    for k in ("rf", "uf", "vf", "wf", "ef"):
        assert np.max(out[k]) < 1e-28, k


def test_friction_reynolds_numpy_equals_loops_and_scales():
    sh = (4, 3, 9)
    r, w = _rand(sh, 7), _rand(sh, 8, -1.0, 1.0)
    for s in (1, 2, 3, 4):
        cF = O.coeffs(s)[0]
        a = D.friction_reynolds_loops(r, w, cF, 0.1, 1.3, 2000.0, viscexp=0.7)
        b = D.friction_reynolds(r, w, cF, 0.1, 1.3, 2000.0, viscexp=0.7)
        assert math.isclose(a, b, rel_tol=1e-13)
        # u_tau ~ sqrt(|dw/dx|): w -> 4 w doubles Re_tau
        assert math.isclose(D.friction_reynolds(r, 4 * w, cF, 0.1, 1.3, 2000.0), 2 * b, rel_tol=1e-13)


def test_friction_reynolds_of_a_linear_profile():
    # w = a x_i on nodes i = 0..: the even mirror makes ub symmetric about the
    # face, so the stencil at node 0 sees w(-1-m) = w(m); for s = 1 the
    # derivative is (w1 - w0) / (2 dx)
    mx, a, dx, Re = 8, 0.3, 0.25, 500.0
    w = np.broadcast_to(a * np.arange(mx, dtype=float), (2, 3, mx)).copy()
    r = np.ones_like(w)
    cF = O.coeffs(1)[0]
    want = math.sqrt((1 / Re) * (a / (2 * dx))) * Re
    assert math.isclose(D.friction_reynolds(r, w, cF, dx, 1.0, Re), want, rel_tol=1e-14)


def _uniform_grid(n, L=2 * math.pi):
    h = L / n
    return dict(x=(np.arange(n) + 0.5) * h, xp=np.ones(n), dxv=np.full(n, h), d_dx=1 / h, d_dy=1 / h, d_dz=1 / h), h


def test_flow_stats_taylor_green_closed_forms():
    n = 16
    g, h = _uniform_grid(n)
    X = g["x"]
    Z, Y, Xg = np.meshgrid(X, X, X, indexing="ij")
    u = np.sin(Xg) * np.cos(Y) * np.cos(Z)
    v = -np.cos(Xg) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    r = np.ones_like(u)
    vol = (2 * math.pi) ** 3
    for s in (1, 2, 3, 4):
        cVF = O.coeffs(s)[0]
        kp = D.modified_wavenumber(cVF, h)
        st = D.flow_stats(r, u, v, w, g, cVF, vol)
        assert math.isclose(st["kinetic"], 1 / 8, rel_tol=1e-13)
        assert math.isclose(st["enstrophy"], 3 * kp * kp / 8, rel_tol=1e-12)
        assert st["dil2"] < 1e-26
        assert math.isclose(st["rho"], 1.0, rel_tol=1e-14)
        # a 1D wave u = sin x: dil = k' cos x
        st = D.flow_stats(r, np.sin(Xg), 0 * u, w, g, cVF, vol)
        assert math.isclose(st["dil2"], kp * kp / 2, rel_tol=1e-12)
        assert st["enstrophy"] < 1e-26
