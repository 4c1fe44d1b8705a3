"""Pins of the oracle's assembled right-hand side (src/cuda_rhs.cu:9-396,
src/calc_stress.cu, src/boundary*.h) by manufactured solutions against the
symbolic continuous equations, free-stream preservation and discrete
conservation."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import mms

GAM, MA, RE, PR, N = 1.4, 0.3, 50.0, 0.7, 0.7


def _max_err(num, ex, mask=None):
    out = []
    for a, b in zip(num, ex):
        d = np.abs(a - b)
        if mask is not None:
            d = d[mask]
        out.append(np.max(d) / max(np.max(np.abs(b)), 1e-300))
    return np.array(out)


def _periodic_case(n, s, fix):
    orc = O.Oracle(mx=n, my=n, mz=n, s=s, sv=s, Re=RE, Pr=PR, Ma=MA, gam=GAM, viscexp=N, fix_y_work=fix)
    g = orc.grid()
    f = mms.periodic_fields()
    orc.set_state(*mms.state_arrays(f, g["x"], g["y"], g["z"], GAM, MA))
    return orc, g, f


_EXACT = {}


def _exact(f, g, key):
    if key not in _EXACT:
        ex = mms.continuous_rhs(*f, gam=GAM, Ma=MA, Re=RE, Pr=PR, n=N)
        _EXACT[key] = mms.lambdify_on_grid(ex, g["x"], g["y"], g["z"])
    return _EXACT[key]


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_mms_periodic_convergence(s):
    """With the corrected y viscous work the discrete RHS converges to the
    continuous one at order 2s in every equation (a dropped term, a wrong
    sign or a wrong index leaves an O(1) error)."""
    errs = []
    sizes = (24, 48)
    for n in sizes:
        orc, g, f = _periodic_case(n, s, fix=1)
        errs.append(_max_err(orc.rhs(), _exact(f, g, ("per", n))))
    order = np.log2(errs[0] / errs[1])
    assert np.all(order > 2 * s - 0.7), (order, errs)


def test_reference_y_viscous_work_term():
    """src/cuda_rhs.cu:175 multiplies the zy stress by dv/dz instead of dw/dy:
    the energy RHS then differs from the corrected one by
    mu (dw/dy + dv/dz)(dv/dz - dw/dy), and nothing else changes."""
    s = 2
    for n in (24, 48):
        ref, g, f = _periodic_case(n, s, fix=0)
        fixd, _, _ = _periodic_case(n, s, fix=1)
        a, b = ref.rhs(), fixd.rhs()
        for q in range(4):
            assert np.array_equal(a[q], b[q])
        rho, u, v, w, T = f
        mu = T ** N / RE
        term = mu * (mms.sp.diff(w, mms.Y) + mms.sp.diff(v, mms.Z)) * (mms.sp.diff(v, mms.Z) - mms.sp.diff(w, mms.Y))
        ex = mms.lambdify_on_grid([term], g["x"], g["y"], g["z"])[0]
        err = np.max(np.abs((a[4] - b[4]) - ex)) / np.max(np.abs(ex))
        if n == 24:
            e0 = err
    assert np.log2(e0 / err) > 2 * s - 0.7


@pytest.mark.parametrize("s", [2, 3])
def test_mms_wall_stretched_convergence(s):
    """Isothermal walls on a tanh-stretched x grid: fields that satisfy the
    ghost rules converge at order 2s away from the walls; the first s cells
    use the mirrored-grid metrics of derivGrid and are only required to
    converge."""
    Lx, Ly, Lz = 2.0, 2.0, 3.0
    errs_in, errs_wall = [], []
    for (mx, my, mz) in ((24, 16, 16), (48, 32, 32)):
        orc = O.Oracle(mx=mx, my=my, mz=mz, s=s, sv=s, periodic_x=0, nonuniform_x=1, stretch=1.8, Lx=Lx, Ly=Ly,
                       Lz=Lz, Re=RE, Pr=PR, Ma=MA, gam=GAM, viscexp=N, fix_y_work=1, twall_bot=1.0, twall_top=1.0)
        g = orc.grid()
        f = mms.wall_fields(Lx, Ly, Lz, Tw=1.0, R=g["Rgas"])
        orc.set_state(*mms.state_arrays(f, g["x"], g["y"], g["z"], GAM, MA))
        ex = mms.lambdify_on_grid(mms.continuous_rhs(*f, gam=GAM, Ma=MA, Re=RE, Pr=PR, n=N), g["x"], g["y"], g["z"])
        num = orc.rhs()
        i = np.arange(mx)
        inner = np.broadcast_to(((i >= s) & (i < mx - s))[None, None, :], (mz, my, mx))
        errs_in.append(_max_err(num, ex, inner))
        errs_wall.append(_max_err(num, ex, ~inner))
    order_in = np.log2(errs_in[0] / errs_in[1])
    order_wall = np.log2(errs_wall[0] / errs_wall[1])
    assert np.all(order_in > 2 * s - 1.0), (order_in, errs_in)
    assert np.all(order_wall > 0.5), (order_wall, errs_wall)


@pytest.mark.parametrize("kw", [dict(), dict(periodic_x=0, nonuniform_x=1, Lx=2.0), dict(s=4, sv=2)])
def test_free_stream_is_exactly_steady(kw):
    cfg = dict(mx=20, my=16, mz=12, s=3, sv=3, viscexp=0.7, Ma=0.5, Re=200.0)
    cfg.update(kw)
    orc = O.Oracle(**cfg)
    shp = (cfg["mz"], cfg["my"], cfg["mx"])
    wall = cfg.get("periodic_x", 1) == 0
    u0 = 0.0 if wall else 0.3  # walls admit only the fluid at rest
    g = orc.grid()
    r = np.full(shp, 1.2)
    T = np.full(shp, 1.0)
    u, v, w = np.full(shp, u0), np.full(shp, -0.2 if not wall else 0.0), np.full(shp, 0.1 if not wall else 0.0)
    e = r * g["Rgas"] * T / 0.4 + 0.5 * r * (u * u + v * v + w * w)
    orc.set_state(r, u, v, w, e)
    for q, a in enumerate(orc.rhs()):
        scale = [1.2, 1.2, 1.2, 1.2, e.max()][q]
        assert np.max(np.abs(a)) <= 1e-13 * scale, q


def test_discrete_conservation():
    """Periodic box: the mass RHS sums to zero; with constant viscosity
    (viscexp = 0) the momentum RHS does too; the energy flux part telescopes."""
    rng = np.random.default_rng(7)
    for s in (1, 2, 3, 4):
        orc = O.Oracle(mx=16, my=12, mz=10, s=s, sv=s, viscexp=0.0, Ma=0.4, Re=100.0)
        g = orc.grid()
        shp = (10, 12, 16)
        r = 1 + 0.1 * rng.uniform(-1, 1, shp)
        u, v, w = (0.3 * rng.uniform(-1, 1, shp) for _ in range(3))
        T = 1 + 0.1 * rng.uniform(-1, 1, shp)
        e = r * g["Rgas"] * T / 0.4 + 0.5 * r * (u * u + v * v + w * w)
        orc.set_state(r, u, v, w, e)
        R = orc.rhs()
        for q in range(4):
            scale = np.sum(np.abs(R[q]))
            assert abs(np.sum(R[q])) < 1e-12 * scale, (s, q)


def test_rhs_symmetry_taylor_green():
    """The Taylor-Green field is even in z; so is every RHS component
    (rho w is odd), on the cell-centred periodic grid."""
    from tests.cases import tgv_fields
    orc = O.Oracle(mx=16, my=16, mz=16, s=3, sv=3)
    f = tgv_fields(orc.grid(), orc.cfg)
    orc.set_state(*f)
    R = orc.rhs()
    for q in (0, 1, 2, 4):
        np.testing.assert_allclose(R[q], R[q][::-1, :, :], atol=1e-12 * np.max(np.abs(R[q])))
    np.testing.assert_allclose(R[3], -R[3][::-1, :, :], atol=1e-12 * max(np.max(np.abs(R[3])), 1e-30))
