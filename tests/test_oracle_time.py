"""Pins of the oracle's time advance and reductions: RK stability
polynomials and temporal order (src/cuda_main.cu:10-11,44-247), CFL time
step (src/calc_stress.cu:140-160), volume integrals and the forcing
controller (src/calc_stress.cu:16-18,98-201, src/cuda_math.cu:143-230)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import channel_fields, tgv_fields


@pytest.mark.parametrize("scheme,poly", [
    ("smr_rk3", [1, 1, 1 / 2, 1 / 6]),
    ("kutta_rk3", [1, 1, 1 / 2, 1 / 6]),
    ("rk4", [1, 1, 1 / 2, 1 / 6, 1 / 24]),
])
def test_stability_polynomial(scheme, poly):
    """One step of y' = lam y is exactly the truncated exponential series."""
    for z in (-2.5, -1.0, -0.3, 0.2, 0.9):
        y1 = O.rk_scalar(scheme, z, 1.0)
        expect = sum(c * z ** k for k, c in enumerate(poly))
        assert y1 == pytest.approx(expect, rel=1e-15, abs=1e-15)


def test_smr_coefficients_consistency():
    """Low-storage RK3 of Spalart-Moser-Rogers: sum of (alpha + beta) equals 1
    (stage times 8/15, 2/3, 1) -- equivalently R(z) matches exp(z) to O(z^4)."""
    eps = 1e-3
    y = O.rk_scalar("smr_rk3", eps, 1.0)
    assert abs(y - np.exp(eps)) < 1e-12


@pytest.mark.parametrize("scheme,order", [("smr_rk3", 3), ("kutta_rk3", 3), ("rk4", 4)])
def test_temporal_order_on_navier_stokes(scheme, order):
    """Fixed grid, dt halved twice at fixed final time: Richardson ratio of
    successive differences is 2^order."""
    states = []
    for k, cfl in enumerate((0.8, 0.4, 0.2)):
        orc = O.Oracle(mx=12, my=12, mz=12, s=2, sv=2, scheme=scheme, cfl=cfl, check_cfl=10 ** 6,
                       check_bulk=10 ** 6, Ma=0.4, Re=50.0)
        orc.set_state(*tgv_fields(orc.grid(), orc.cfg))
        orc.step(4 * 2 ** k)
        states.append(orc.get_state())
    d1 = np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(states[0], states[1])))
    d2 = np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(states[1], states[2])))
    p = np.log2(d1 / d2)
    assert abs(p - order) < 0.3, p


def test_cfl_time_step_closed_form():
    """Uniform state: dt = CFL / max((|u|+c)/dx, (|v|+c)/dy, (|w|+c)/dz,
    mu/dx^2, ...), c = sqrt(gam (gam-1) e_int)."""
    cfg = dict(mx=16, my=12, mz=10, s=2, sv=2, Lx=2.0, Ly=3.0, Lz=4.0, Ma=0.5, Re=2.0, cfl=0.6, viscexp=1.0)
    orc = O.Oracle(**cfg)
    g = orc.grid()
    shp = (10, 12, 16)
    r, T = 1.1, 1.3
    u, v, w = 0.4, -0.7, 0.2
    e = r * g["Rgas"] * T / 0.4 + 0.5 * r * (u * u + v * v + w * w)
    orc.set_state(np.full(shp, r), np.full(shp, u), np.full(shp, v), np.full(shp, w), np.full(shp, e))
    c = np.sqrt(1.4 * 0.4 * g["Rgas"] * T / 0.4)
    dx, dy, dz = 2.0 / 16, 3.0 / 12, 4.0 / 10
    mu = T / 2.0
    inv = max((abs(u) + c) / dx, (abs(v) + c) / dy, (abs(w) + c) / dz, mu / dx ** 2, mu / dy ** 2, mu / dz ** 2)
    assert orc.calc_dt() == pytest.approx(0.6 / inv, rel=1e-13)


def test_volume_average_taylor_green():
    """<u.u> of the Taylor-Green vortex on the cell-centred grid is exactly 1/4
    (discrete orthogonality of the sampled modes)."""
    orc = O.Oracle(mx=16, my=16, mz=16, s=2, sv=2)
    orc.set_state(*tgv_fields(orc.grid(), orc.cfg))
    par1, par2 = orc.bulk()
    assert par1 == pytest.approx(0.25, rel=1e-14)
    assert np.isnan(par2)


def test_volume_integral_stretched_grid():
    orc = O.Oracle(mx=24, my=8, mz=8, s=3, sv=3, periodic_x=0, nonuniform_x=1, Lx=2.0, Ly=3.0, Lz=5.0)
    assert orc.vol_integral(np.ones(24 * 8 * 8)) == pytest.approx(2.0 * 3.0 * 5.0, rel=1e-13)


def test_forcing_controller_closed_form():
    """dpdz_new = 0.99 dpdz - 0.5 (int(rho w)/int(rho) - 1), evaluated at the
    start of the step (src/calc_stress.cu:16-18)."""
    cfg = dict(mx=16, my=8, mz=8, s=2, sv=2, periodic_x=0, nonuniform_x=1, forcing=1, Lx=2.0, Ly=2.0, Lz=4.0,
               Ma=1.5, Re=3000.0, viscexp=0.7, dpdz0=0.00372, check_cfl=1)
    orc = O.Oracle(**cfg)
    g = orc.grid()
    f = channel_fields(g, orc.cfg, seed=3)
    orc.set_state(*f)
    rw = f[0] * f[3]
    wgt = g["dxv"][None, None, :] * (2.0 / 8) * (4.0 / 8)
    ub = np.sum(rw * wgt) / np.sum(f[0] * wgt)
    orc.step(1)
    assert orc.dpdz == pytest.approx(0.99 * 0.00372 - 0.5 * (ub - 1.0), rel=1e-12)
    # bulk statistic of the forced case: int(rho w)/int(rho) and int(rhoE)
    orc2 = O.Oracle(**cfg)
    orc2.set_state(*f)
    p1, p2 = orc2.bulk()
    assert p1 == pytest.approx(ub, rel=1e-13)
    assert p2 == pytest.approx(np.sum(f[4] * wgt), rel=1e-13)


def test_time_array_accumulates_dt():
    orc = O.Oracle(mx=12, my=12, mz=12, s=2, sv=2, check_cfl=2, check_bulk=3)
    orc.set_state(*tgv_fields(orc.grid(), orc.cfg))
    p1, p2, t = orc.step(7)
    assert np.all(np.diff(t) > 0)
    assert t[-1] == pytest.approx(orc.time)
    assert np.array_equal(np.isnan(p1), np.array([k % 3 != 0 for k in range(7)]))
