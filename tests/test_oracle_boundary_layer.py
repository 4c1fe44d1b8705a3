"""Independent pins of the oracle's boundary-layer rules: the half-tanh wall
grid (src/init.cpp:41-50), the ghost fills of src/boundary.h:150-180 as
src/boundary_condition_x.h and src/boundary_condition_z.h apply them, and
the wall-normal blowing of src/perturbation.h:25-53.

None of these restates a ghost formula.  Each rule is pinned by what it must
do to fields the discrete operators treat exactly:
  - x top and both z ends extrapolate linearly about the end node, so a
    field linear in the grid index keeps its exact central derivative up to
    the last node (a periodic wrap or a mirror would not);
  - at the bottom x boundary the velocities are odd about the face, so
    v = c (i + 1/2) is linear through the ghosts as well;
  - p and T are even about the bottom face (adiabatic): a uniform state at
    rest has a zero right-hand side whatever the wall temperatures say (the
    isothermal rule heats it);
  - pressure is extrapolated on its own in z and at the x top, not formed
    from extrapolated rho and T: a uniform pressure over varying rho keeps
    every momentum right-hand side zero at rest;
  - the blowing lives on k in [90, 130], vanishes at both ends and at 110,
    is odd about 110, is a cosine of period my in j, takes its hand-evaluated
    value at mid-band, reaches every one of the S ghosts, and is evaluated
    at the time after the step's advance.
"""
import numpy as np
import pytest

from oracle import oracle as O

SV_SUM = {1: 0.5, 2: 2.0 / 3.0 - 1.0 / 12.0, 3: 3.0 / 4.0 - 3.0 / 20.0 + 1.0 / 60.0}


def _bl(**kw):
    base = dict(mx=24, my=8, mz=20, s=3, sv=2, periodic_x=0, nonuniform_x=1, boundary_layer=1, perturbed=0,
                stretch=5.0, Lx=20.0, Ly=7.0, Lz=50.0, Re=1500.0, Pr=0.75, Ma=0.35, viscexp=1.5,
                twall_bot=1.0, twall_top=1.0)
    base.update(kw)
    return O.Oracle(**base)


def _state(o, r, u, v, w, p):
    gam = o.cfg.gam
    sh = (o.cfg.mz, o.cfg.my, o.cfg.mx)
    r, u, v, w, p = [np.broadcast_to(np.asarray(a, dtype=np.float64), sh).copy() for a in (r, u, v, w, p)]
    e = p / (gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    o.set_state(r, u, v, w, e)
    return r, u, v, w, e


def _index_grids(o):
    c = o.cfg
    k, j, i = np.meshgrid(np.arange(c.mz), np.arange(c.my), np.arange(c.mx), indexing="ij")
    return i.astype(float), j.astype(float), k.astype(float)


def test_wall_grid_is_the_lower_half_of_the_tanh_profile():
    o = _bl(mx=64, nonuniform_x=1)
    g = o.grid()
    x = g["x"]
    dx = np.diff(x)
    assert x[0] > 0 and x[-1] < o.cfg.Lx
    assert np.all(dx > 0) and np.all(np.diff(dx) > 0), "spacing grows away from the wall x = 0"
    # the channel grid (both walls clustered) is symmetric about Lx/2; this one is not
    ch = O.Oracle(mx=64, my=8, mz=20, s=3, sv=2, periodic_x=0, nonuniform_x=1, stretch=5.0, Lx=20.0)
    xc = ch.grid()["x"]
    assert np.allclose(xc + xc[::-1], 20.0, rtol=0, atol=1e-12)
    assert dx[-1] / dx[0] > 10.0
    # faces at 0 and Lx: the first and last cell centres sit half a cell in
    # from them to first order in the spacing
    assert abs(x[0] - dx[0] / 2) < 0.25 * dx[0]
    assert abs(o.cfg.Lx - x[-1] - dx[-1] / 2) < 0.25 * dx[-1]


@pytest.mark.parametrize("nonuni", [0, 1])
@pytest.mark.parametrize("s,sv", [(3, 2), (2, 2), (4, 3)])
def test_linear_fields_keep_exact_derivatives_at_the_ends(nonuni, s, sv):
    o = _bl(s=s, sv=sv, nonuniform_x=nonuni)
    g = o.grid()
    i, j, k = _index_grids(o)
    cu, cv, cw = 1e-3, -2e-3, 3e-3
    # v odd about the bottom face and linear; u, w linear in k (z ends)
    _state(o, 1.0, 0.1 + cu * k, cv * (i + 0.5), 0.2 + cw * k, 1.0 / (o.cfg.gam * o.cfg.Ma ** 2))
    o.rhs()
    d_dx, d_dz = 1.0 / g["dx"], 1.0 / (g["z"][1] - g["z"][0])
    xp = g["xp"][None, None, :] if nonuni else 1.0  # the metric applies on the stretched grid only
    np.testing.assert_allclose(o.field("dvdx"), cv * d_dx * xp * np.ones_like(i), rtol=1e-11, atol=0)
    np.testing.assert_allclose(o.field("dudz"), cu * d_dz * np.ones_like(i), rtol=1e-11, atol=0)
    np.testing.assert_allclose(o.field("dwdz"), cw * d_dz * np.ones_like(i), rtol=1e-11, atol=0)
    # a periodic wrap in z breaks the end derivatives of the same field
    p = O.Oracle(**{f: getattr(o.cfg, f) for f, _ in O.OrcConfig._fields_ if f != "boundary_layer"},
                 boundary_layer=0)
    _state(p, 1.0, 0.1 + cu * k, cv * (i + 0.5), 0.2 + cw * k, 1.0 / (o.cfg.gam * o.cfg.Ma ** 2))
    p.rhs()
    assert abs(p.field("dwdz")[0, 0, 0] - cw * d_dz) > 1e-3 * abs(cw * d_dz)


def test_u_top_extrapolation_in_x():
    # u linear in the index near the top: exact derivative at the last node
    o = _bl(nonuniform_x=0)
    g = o.grid()
    i, _, _ = _index_grids(o)
    c = 2e-3
    _state(o, 1.0, c * i, 0.0, 0.0, 1.0 / (o.cfg.gam * o.cfg.Ma ** 2))
    o.rhs()
    d = o.field("dudx")
    mx, sv = o.cfg.mx, o.cfg.sv
    np.testing.assert_allclose(d[:, :, sv:], c / g["dx"], rtol=1e-11)
    # at the bottom u is odd about the face, so c*i is not linear through the
    # ghosts: the first node sees u(-d) = -u(d-1), not the linear continuation
    assert abs(d[0, 0, 0] - c / g["dx"]) > 1e-3 * c / g["dx"]
    assert mx > 2 * sv


def test_rest_state_is_adiabatic_at_the_bottom():
    p0 = 1.0 / (1.4 * 0.35 ** 2)
    for tb, tt in ((1.5, 0.8), (1.0, 1.0)):
        o = _bl(twall_bot=tb, twall_top=tt)
        _state(o, 1.0, 0.0, 0.0, 0.0, p0)
        for a in o.rhs():
            assert np.max(np.abs(a)) <= 1e-12 * p0
    # the isothermal wall with Tw != T heats the same state
    w = O.Oracle(mx=24, my=8, mz=20, s=3, sv=2, periodic_x=0, nonuniform_x=1, stretch=5.0, Lx=20.0, Ly=7.0,
                 Lz=50.0, Re=1500.0, Pr=0.75, Ma=0.35, viscexp=1.5, twall_bot=1.5, twall_top=0.8)
    _state(w, 1.0, 0.0, 0.0, 0.0, p0)
    assert np.max(np.abs(w.rhs()[4])) > 1e-6


def test_pressure_is_extrapolated_on_its_own():
    o = _bl()
    i, j, k = _index_grids(o)
    p0 = 1.0 / (o.cfg.gam * o.cfg.Ma ** 2)
    # rho varies (nonlinearly) in x and z, T = p0/(rho R): p is uniform
    rho = 1.0 + 0.3 * (k / o.cfg.mz) ** 2 + 0.2 * np.sin(0.4 * i)
    _state(o, rho, 0.0, 0.0, 0.0, p0)
    rr, ru, rv, rw, re = o.rhs()
    assert np.max(np.abs(rr)) == 0.0
    for a in (ru, rv, rw):
        assert np.max(np.abs(a)) <= 1e-11 * p0
    # heat conduction acts on the varying T
    assert np.max(np.abs(re)) > 1e-6


def test_blowing_band_zeros_and_antisymmetry():
    o = _bl(my=16, perturbed=1)
    t = 3.7
    for kk in (0, 89, 131, 500):
        assert o.perturbation(3, kk, t) is None
    for kk in (90, 110, 130):
        assert o.perturbation(3, kk, t) == 0.0
    for m in range(1, 20):
        for jj in (0, 5):
            a, b = o.perturbation(jj, 110 - m, t), o.perturbation(jj, 110 + m, t)
            assert a != 0.0 and b == -a


def test_blowing_spanwise_cosine_and_mid_band_value():
    my = 16
    o = _bl(my=my, perturbed=1)
    t = 2.3
    vals = np.array([o.perturbation(jj, 100, t) for jj in range(my)])
    c = np.cos(2 * np.pi * np.arange(my) / my)
    A = vals.mean()
    B = 2 * np.dot(vals - A, c) / my
    np.testing.assert_allclose(vals, A + B * c, rtol=0, atol=1e-15)
    # mid-band xi = 1/2: g = 15.1875/32 - 35.4375/16 + 20.25/8 = 0.791015625
    g = 0.791015625
    np.testing.assert_allclose(A, 2e-4 * g * np.sin(1500.0 * 121e-6 * t), rtol=1e-14)
    np.testing.assert_allclose(B, 3e-5 * g * np.sin(10.0 * t), rtol=1e-12)
    # at a time where sin(10 t) = 0 the spanwise part vanishes
    t2 = np.pi
    v2 = np.array([o.perturbation(jj, 100, t2) for jj in range(my)])
    assert np.ptp(v2) < 1e-18


@pytest.mark.parametrize("sv", [1, 2, 3])
def test_blowing_fills_every_ghost(sv):
    o = _bl(mz=140, s=3, sv=sv, perturbed=1, nonuniform_x=1)
    g = o.grid()
    p0 = 1.0 / (o.cfg.gam * o.cfg.Ma ** 2)
    _state(o, 1.0, 0.0, 0.0, 0.0, p0)
    o.time = 2.3
    o.rhs()
    d = o.field("dudx")
    for kk in (95, 100, 125):
        for jj in (0, 3):
            val = o.perturbation(jj, kk, 2.3)
            expect = -val * SV_SUM[sv] * g["xp"][0] / g["dx"]  # stretched grid: metric xp
            np.testing.assert_allclose(d[kk, jj, 0], expect, rtol=1e-12)
    assert np.all(d[:90] == 0.0) and np.all(d[131:] == 0.0)
    assert np.all(d[:, :, 2 * o.cfg.s:] == 0.0)


def test_blowing_uses_the_time_after_the_advance():
    # from t = 0 the blowing is zero at the step's start; the state moves only
    # if the stages see t = dt
    o = _bl(mz=140, perturbed=1, scheme=0)
    p0 = 1.0 / (o.cfg.gam * o.cfg.Ma ** 2)
    st = _state(o, 1.0, 0.0, 0.0, 0.0, p0)
    o.step(1)
    assert o.time > 0
    r1, u1, v1, w1, e1 = o.get_state()
    # three stages carry it at most 3 s planes out of the band
    assert np.max(np.abs(u1[90:131])) > 0 and np.max(np.abs(u1[:90 - 3 * o.cfg.s])) == 0.0
    # without the blowing the rest state stays at rest
    q = _bl(mz=140, perturbed=0, scheme=0)
    _state(q, 1.0, 0.0, 0.0, 0.0, p0)
    q.step(1)
    assert np.max(np.abs(q.get_state()[1])) == 0.0
    assert st is not None
