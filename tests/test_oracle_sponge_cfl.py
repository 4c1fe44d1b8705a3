"""Independent pins of the oracle's sponge term (src/sponge.cu:31-41) and of
the convective branch of its CFL time step (src/calc_stress.cu:140-160).

The sponge adds (sigma_x(i) + sigma_z(k)) (q_ref(i, k) - q) to each
conserved right-hand side.  It is pinned without restating that formula:
  - a state equal to its reference leaves the right-hand side bitwise
    unchanged (catches a reference compared with the wrong variable, e.g.
    u instead of rho u);
  - a reference offset by a constant dq per field adds exactly
    (sigma_x(i) + sigma_z(k)) dq, the same for every j (catches a sign
    error, a swapped i/k index, a dropped sigma or a reference indexed by j);
  - with sigma_z = 0 the added term is independent of k, with sigma_x = 0
    independent of i (separability).
The CFL step is pinned on closed forms where the convective limit wins
(high Reynolds number): uniform states on uniform and stretched grids, and a
single fast point that sets the minimum.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import random_smooth_fields


def _xz_state(orc, seed):
    """A smooth state that varies in x and z only, so that it can equal a
    sponge reference (which is a function of (i, k))."""
    g = orc.grid()
    f = random_smooth_fields(g, orc.cfg, seed=seed)
    return [np.broadcast_to(a[:, :1, :], a.shape).copy() for a in f]


def _pair(**kw):
    base = dict(mx=20, my=6, mz=14, s=3, sv=2, Lx=2.0, Ly=1.5, Lz=3.0, Re=300.0, Ma=0.3)
    base.update(kw)
    on = O.Oracle(sponge=1, **base)
    off = O.Oracle(sponge=0, **base)
    return on, off


def _refs_of(state, shift=(0, 0, 0, 0, 0)):
    r, u, v, w, e = state
    cons = [r, r * u, r * v, r * w, e]
    # (i, k) reference from the j = 0 row, as [k][i] -> stored i + k*mx
    return [np.ascontiguousarray(c[:, 0, :] + d) for c, d in zip(cons, shift)]


@pytest.mark.parametrize("periodic_x", [1, 0])
def test_sponge_vanishes_on_its_reference_state(periodic_x):
    on, off = _pair(periodic_x=periodic_x, nonuniform_x=1 - periodic_x)
    st = _xz_state(on, seed=11)
    rng = np.random.default_rng(3)
    sx, sz = rng.uniform(0.5, 2.0, on.cfg.mx), rng.uniform(0.5, 2.0, on.cfg.mz)
    on.set_sponge(sx, sz, _refs_of(st))
    on.set_state(*st)
    off.set_state(*st)
    for a, b in zip(on.rhs(), off.rhs()):
        assert np.array_equal(a, b)


def test_sponge_offset_reference_adds_sigma_times_offset():
    on, off = _pair()
    st = _xz_state(on, seed=5)
    mx, my, mz = on.cfg.mx, on.cfg.my, on.cfg.mz
    rng = np.random.default_rng(7)
    sx, sz = rng.uniform(0.1, 3.0, mx), rng.uniform(0.1, 3.0, mz)
    dq = (0.125, -0.25, 0.375, 0.0625, -0.5)
    on.set_sponge(sx, sz, _refs_of(st, dq))
    on.set_state(*st)
    off.set_state(*st)
    sig = sz[:, None, None] + sx[None, None, :]
    for q, (a, b) in enumerate(zip(on.rhs(), off.rhs())):
        expect = np.broadcast_to(sig * dq[q], (mz, my, mx))
        scale = np.max(np.abs(b)) + 1.0
        assert np.max(np.abs((a - b) - expect)) <= 1e-13 * scale * 8, q


@pytest.mark.parametrize("which", ["x", "z"])
def test_sponge_is_separable(which):
    on, off = _pair()
    st = _xz_state(on, seed=9)
    mx, mz = on.cfg.mx, on.cfg.mz
    rng = np.random.default_rng(13)
    sx = rng.uniform(0.1, 3.0, mx) if which == "x" else np.zeros(mx)
    sz = rng.uniform(0.1, 3.0, mz) if which == "z" else np.zeros(mz)
    dq = (1.0, 1.0, 1.0, 1.0, 1.0)
    on.set_sponge(sx, sz, _refs_of(st, dq))
    on.set_state(*st)
    off.set_state(*st)
    d = on.rhs()[0] - off.rhs()[0]
    # the added term depends on i only (x sponge) or on k only (z sponge)
    axis_const = (0, 1) if which == "x" else (1, 2)
    spread = np.max(d, axis=axis_const) - np.min(d, axis=axis_const)
    scale = np.max(np.abs(off.rhs()[0])) + 1.0
    assert np.max(spread) <= 1e-13 * scale * 8
    prof = d.mean(axis=axis_const)
    np.testing.assert_allclose(prof, sx if which == "x" else sz, rtol=0, atol=1e-13 * scale * 8)


def _uniform(orc, r, T, u, v, w):
    g = orc.grid()
    c = orc.cfg
    shp = (c.mz, c.my, c.mx)
    e = r * g["Rgas"] * T / (c.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    return [np.full(shp, r), np.full(shp, u), np.full(shp, v), np.full(shp, w), np.full(shp, e)], g


@pytest.mark.parametrize("stretched", [False, True])
def test_cfl_convective_limit_closed_form(stretched):
    """High Re: dt = CFL min_i dx_i / (|u| + c) etc., c = sqrt(gam R T)
    (sound speed of the perfect gas, independent of the code's
    gam (gam-1) e_int form)."""
    kw = dict(mx=18, my=12, mz=10, s=2, sv=2, Lx=2.0, Ly=3.0, Lz=4.0, Ma=0.8, Re=1.0e6, cfl=0.7)
    if stretched:
        kw.update(periodic_x=0, nonuniform_x=1, stretch=1.8)
    orc = O.Oracle(**kw)
    u, v, w, T = 0.9, -0.4, 0.3, 1.2
    st, g = _uniform(orc, 1.05, T, u, v, w)
    orc.set_state(*st)
    c = np.sqrt(1.4 * g["Rgas"] * T)
    dxmin = float(np.min(g["dxv"]))
    dy, dz = 3.0 / 12, 4.0 / 10
    conv = max((abs(u) + c) / dxmin, (abs(v) + c) / dy, (abs(w) + c) / dz)
    visc = (T / 1.0e6) / dxmin ** 2
    assert conv > 1e3 * visc  # the convective limit wins by far
    assert orc.calc_dt() == pytest.approx(0.7 / conv, rel=1e-13)
    if stretched:
        # the smallest cell is at the wall, where the tanh grid clusters
        assert int(np.argmin(g["dxv"])) in (0, 17) and dxmin < 0.8 * 2.0 / 18


def test_cfl_single_fast_point_sets_the_step():
    """One point with a large velocity along y sets dt = CFL dy / (|v| + c)."""
    orc = O.Oracle(mx=12, my=16, mz=8, s=2, sv=2, Lx=1.0, Ly=2.0, Lz=1.0, Ma=0.5, Re=1.0e6, cfl=0.5)
    st, g = _uniform(orc, 1.0, 1.0, 0.0, 0.0, 0.0)
    r, u, v, w, e = st
    vfast = 7.5
    v[3, 11, 5] = -vfast
    e[3, 11, 5] += 0.5 * r[3, 11, 5] * vfast ** 2  # same internal energy, so the same c
    orc.set_state(r, u, v, w, e)
    c = np.sqrt(1.4 * g["Rgas"] * 1.0)
    dx, dy = 1.0 / 12, 2.0 / 16
    expect = 0.5 / max((vfast + c) / dy, c / dx)
    assert orc.calc_dt() == pytest.approx(expect, rel=1e-13)
