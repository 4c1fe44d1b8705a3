"""Boundary layer on the GPU against the FP64 oracle: the reference's
default case (src/globals.h: boundaryLayer and perturbed, stretch 5, order
6 advective / 4 viscous, sponge, viscexp 1.5, Re 1500, Ma 0.35, low-storage
RK3) at a reduced size, 48 x 16 x 160 with the blowing band (z planes
90..130) inside the domain, plus other stencil orders, schemes, viscosity
laws and grids.

Each case starts from the sponge's Blasius reference state with smooth
noise (so every direction has gradients), at simulation time 2.3 (so the
blowing is on from the first stage); it compares one right-hand side and a
few steps.  Tolerances: relative L2 <= 1e-12, Linf <= 1e-10 per conserved
variable (momenta normalised jointly); time and dt relative 1e-12.
"""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from cudanavierstokes_amd import inflow
from tests import cases
from tests.helpers import assert_close, flush_print, oracle_from, state_errors

pytestmark = pytest.mark.gpu

TOL = (1e-12, 1e-10)
T0 = 2.3


def _state(cfg, g, refs, seed=3, noise=0.02):
    """The reference state (constant along y) with smooth noise."""
    sh = (cfg.mz, cfg.my, cfg.mx)
    r0 = np.broadcast_to(refs[0][:, None, :], sh)
    u0 = np.broadcast_to((refs[1] / refs[0])[:, None, :], sh)
    w0 = np.broadcast_to((refs[3] / refs[0])[:, None, :], sh)
    Rgas = 1.0 / (cfg.gam * cfg.Ma ** 2)
    nz = cases.random_smooth_fields(g, cfg, seed=seed, amp=1.0)
    r = r0 * (1.0 + noise * (nz[0] - 1.0))
    u = u0 + noise * nz[1]
    v = noise * nz[2]
    w = w0 + noise * nz[3]
    p = Rgas * (1.0 + noise * np.tanh(nz[1] * nz[2]))
    e = p / (cfg.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    return [np.ascontiguousarray(a) for a in (r, u, v, w, e)]


def _pair(cfg, prof):
    s = cns.Solver(cfg)
    g = s.grid()
    sx, sz = inflow.sponge_strengths(g["x"], g["z"], cfg.Lx, cfg.Lz)
    refs = inflow.reference_fields(g["x"], g["z"], prof, cfg.Re, cfg.gam, cfg.Ma)
    orc = oracle_from(cfg)
    if cfg.sponge:
        s.set_sponge(sx, sz, refs)
        orc.set_sponge(sx, sz, refs)
    f = _state(cfg, orc.grid(), refs)
    s.set_state(*f)
    orc.set_state(*f)
    ctl = s.control()
    ctl["time"] = T0
    s.set_control(ctl)
    orc.time = T0
    return s, orc


@pytest.fixture(scope="module")
def prof():
    return inflow.similarity_profile(gam=1.4, Ma=0.35, Pr=0.75, n=300, n_out=400)


CASES = [
    ("default-reduced", dict(mx=48, my=16, mz=160, check_cfl=2, check_bulk=2), True),
    ("o8-rk4-unitv-uniform", dict(mx=32, my=16, mz=144, order=8, visc_order=8, scheme="rk4", viscexp=1.0,
                                  nonuniform_x=0, sponge=0, check_cfl=1), False),
    ("o4-kutta-noblow", dict(mx=16, my=32, mz=50, order=4, visc_order=2, scheme="kutta_rk3", perturbed=0,
                             sponge=0, check_cfl=3), False),
    ("o2-rk3-short-chunks", dict(mx=32, my=16, mz=37, order=2, visc_order=2, check_cfl=1), True),
]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,kw,tune", CASES, ids=[c[0] for c in CASES])
def test_boundary_layer_against_oracle(name, kw, tune, prof):
    cfg = cns.boundary_layer(**kw)
    flush_print(f"[case] boundary layer {name}: {cfg}")
    s, orc = _pair(cfg, prof)
    try:
        e = cases.rel_errors(s.rhs(), orc.rhs())
        flush_print(f"{name} rhs errors {e}")
        assert_close(e, *TOL, f"{name} rhs")
        pencil = s.tune() if tune else None
        nsteps = 5
        _, _, t = s.step(nsteps)
        _, _, to = orc.step(nsteps)
        e = state_errors(s.get_state(), orc.get_state())
        flush_print(f"{name} pencil={pencil} state errors after {nsteps} steps {e}")
        assert_close(e, *TOL, f"{name} state")
        np.testing.assert_allclose(t, to, rtol=1e-12, atol=0)
        assert s.scalars()["dt"] == pytest.approx(orc.dt, rel=1e-12)
    finally:
        s.close()
        orc.close()


def test_boundary_layer_grid_matches_oracle():
    cfg = cns.boundary_layer(mx=48, my=16, mz=160)
    s = cns.Solver(cfg)
    try:
        g, go = s.grid(), oracle_from(cfg).grid()
        for k in ("x", "xp", "xpp", "dxv"):
            np.testing.assert_allclose(g[k], go[k], rtol=1e-14, atol=1e-14)
        assert g["x"][0] < 0.1 and np.all(np.diff(np.diff(g["x"])) > 0)
    finally:
        s.close()


@pytest.mark.parametrize("kw,what", [
    (dict(precision=1), "FP64 storage only"),
    (dict(mx=40), "multiple of 16"),
    (dict(periodic_x=1), "periodic_x = 0"),
    (dict(mz=6, order=6), "nz > 2*stencil_size"),
])
def test_boundary_layer_rejected_configurations(kw, what):
    cfg = cns.boundary_layer(**dict(dict(mx=48, my=16, mz=64), **kw))
    with pytest.raises(cns.CnsError, match=what.replace("*", r"\*")):
        cns.Solver(cfg)


def test_boundary_layer_not_batched(prof):
    cfg = cns.boundary_layer(mx=16, my=16, mz=32, sponge=0)
    a, orc = _pair(cfg, prof)
    orc.close()
    try:
        with pytest.raises(cns.CnsError, match="boundary-layer solvers run alone"):
            cns.Batch([a])
    finally:
        a.close()


@pytest.mark.timeout(300)
def test_default_boundary_layer_full_size_runs_stably():
    """The reference's default case at its own size (240 x 64 x 2048) from
    the Blasius inflow state: 100 steps stay finite, the density within a few
    percent of the free stream, the mass within 1e-5 of the start, the
    wall-normal velocity next to the wall at the blowing's scale (its amplitude
    is 2e-4 times a profile below 1), and the blowing's spanwise pattern only
    inside its band."""
    cfg = cns.boundary_layer()
    flush_print(f"[case] boundary layer full size {cfg.mx}x{cfg.my}x{cfg.mz}")
    s = cns.Solver(cfg)
    try:
        inflow.calculate_sponge(s, set_state=True)
        m0 = float(np.sum(s.get_state()[0]))
        s.step(100, outputs=False)
        r, u, v, w, e = s.get_state()
        for a in (r, u, v, w, e):
            assert np.isfinite(a).all()
        assert 0.9 < r.min() and r.max() < 1.1
        assert abs(float(np.sum(r)) / m0 - 1.0) < 1e-5
        assert np.abs(u[:, :, 0]).max() < 2e-4
        # the inflow state is uniform along y; only the blowing (through its
        # spanwise cosine) makes the first node's u vary with y, in its band
        assert np.std(u[100:121, :, 0], axis=1).max() > 1e-9
        assert np.std(u[:60, :, 0], axis=1).max() < 1e-14
    finally:
        s.close()
