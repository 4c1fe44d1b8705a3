"""Boundary-layer inflow data on the host: sponge zones, the natural cubic
spline, the compressible similarity profile (pinned by the incompressible
Blasius constant and the Crocco-Busemann relation) and the sponge reference
state (src/sponge.cu:83-313, python-utils/selfSimilarSol.py)."""
import math

import numpy as np
import pytest

from cudanavierstokes_amd import inflow as fl


def test_sponge_zone_closed_forms():
    x = np.array([0.0, 2.9, 3.0, 3.5, 4.0])
    z = np.array([0.0, 10.0, 20.0, 30.0, 40.0, 50.0, 60.0])
    sx, sz = fl.sponge_strengths(x, z, Lx=4.0, Lz=60.0)
    assert np.allclose(sx, [0, 0, 0, 0.25, 1.0])                       # 1 * ((x-3)/1)^2
    assert np.allclose(sz, [0.5, 0.125, 0.0, 0, 0.0, 0.125, 0.5])     # inlet 0.5((20-z)/20)^2, outlet mirror
    sx, sz = fl.sponge_strengths(x, z, 4.0, 60.0, top=(2.0, 0.0, 2.0), inlet=(1.0, 30.0, 1.0))
    assert not sx.any()
    assert np.isclose(sz[1], 1.0 * (20.0 / 30.0)) and np.isclose(sz[-1], 0.5)


def test_natural_spline_matches_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(3)
    xk = np.cumsum(rng.uniform(0.1, 1.0, 40))
    yk = np.sin(xk) + rng.normal(0, 0.1, 40)
    ev = fl.natural_spline(xk, yk)
    ref = CubicSpline(xk, yk, bc_type="natural")
    xs = np.linspace(xk[0] - 0.5, xk[-1] + 0.5, 777)
    assert np.allclose(ev(xs), ref(xs), rtol=1e-11, atol=1e-11)
    assert np.allclose(ev(xk), yk, rtol=0, atol=1e-13)
    with pytest.raises(ValueError):
        fl.natural_spline([0.0, 1.0, 1.0], [0.0, 1.0, 2.0])


def test_similarity_profile_incompressible_limit_is_blasius():
    p = fl.similarity_profile(Ma=1e-3, Pr=0.72)
    # f''' + f f'' = 0: f''(0) = 0.469600 (Blasius)
    assert abs(p["wall_shear"] - 0.469600) < 2e-6
    assert abs(p["wall_temperature"] - 1.0) < 1e-6
    i = np.argmin(abs(p["y"] - 1.0))
    assert abs(p["y"][i] - 1.0) < 1e-12 and p["w"][i] > 0.99
    assert p["y"].size == 1000 and np.all(np.diff(p["y"]) > 0)


def test_similarity_profile_crocco_busemann():
    gam, Ma = 1.4, 0.8
    p = fl.similarity_profile(gam=gam, Ma=Ma, Pr=1.0, exp_mu=1.0)
    Ec = (gam - 1) * Ma * Ma
    Rgas = 1.0 / (gam * Ma * Ma)
    T = p["e"] * (gam - 1) / Rgas
    # Pr = 1, adiabatic wall: T = 1 + Ec/2 (1 - U^2)
    assert np.max(np.abs(T - (1 + 0.5 * Ec * (1 - p["w"] ** 2)))) < 1e-6
    assert abs(p["wall_temperature"] - (1 + 0.5 * Ec)) < 1e-6
    assert np.allclose(p["r"], 1.0 / T, rtol=1e-12)


def test_reference_fields():
    p = fl.similarity_profile(Ma=0.3)
    x = np.linspace(0.05, 3.0, 12)
    z = np.array([0.0, 30.0, 300.0])
    Re, gam, Ma = 300.0, 1.4, 0.3
    r, ru, rv, rw, e = fl.reference_fields(x, z, p, Re, gam, Ma)
    ev = {k: fl.natural_spline(p["y"], p[k]) for k in ("r", "u", "w")}
    assert np.allclose(r[0], ev["r"](x), rtol=1e-14)
    assert np.allclose(rw[0], ev["r"](x) * ev["w"](x), rtol=1e-14)
    s = math.sqrt(1 + 300.0 / Re)
    assert np.allclose(r[2], ev["r"](x / s), rtol=1e-14)
    assert np.allclose(ru[2] / r[2], ev["u"](x / s) / (s * Re), rtol=1e-13)
    assert not rv.any()
    Rgas = 1.0 / (gam * Ma * Ma)
    assert np.allclose(e - 0.5 * (ru ** 2 + rw ** 2) / r, Rgas / (gam - 1), rtol=1e-13)
