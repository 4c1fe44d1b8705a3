"""Batched advance (cns_batch_*) against each member advanced on its own.

A batch runs every member through the arithmetic of cns_step (same tiles,
same operators, same step control).  The batched kernel is compiled
separately from the single-instance one and the compiler fuses a few
multiply-adds differently, so results agree to rounding rather than bitwise:
relative L2 <= 1e-12 and Linf <= 1e-10 per conserved variable for FP64
storage, 1e-6 / 1e-5 for FP32 storage (scalars dt, dp/dz, time: relative
1e-12 / 1e-6).  Members of different sizes, schemes, viscosity laws and
boundary kinds form several groups; the steps are split over several
cns_batch_step calls so that the captured graph is replayed with host work
in between.
"""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests.helpers import assert_close, flush_print, initial_fields, oracle_from, state_errors

pytestmark = pytest.mark.gpu

TOL = {0: (1e-12, 1e-10, 1e-12), 1: (1e-6, 1e-5, 1e-6)}  # precision -> L2, Linf, scalars


def _same(name, precision, got, want, sc_got=None, sc_want=None):
    l2, linf, rs = TOL[precision]
    errs = state_errors(got, want)
    assert_close(errs, l2, linf, name)
    if sc_got is not None:
        for k in ("dt", "dpdz", "time"):
            assert sc_got[k] == pytest.approx(sc_want[k], rel=rs, abs=1e-300), (name, k)


CHANNEL = dict(mx=24, my=16, mz=20, stencil_size=3, stencil_visc=3, periodic_x=0, nonuniform_x=1, forcing=1,
               Lx=2.0, Ly=4 * np.pi / 3, Lz=4 * np.pi, Re=3000.0, Pr=0.72, Ma=1.5, viscexp=0.7, scheme=0,
               check_cfl=3)

MEMBERS = [
    ("tgv16_rk4", "tgv", dict(mx=16, my=16, mz=16, stencil_size=2, stencil_visc=2, scheme=2, check_cfl=3)),
    ("box_rk4", "tgv", dict(mx=24, my=20, mz=18, stencil_size=2, stencil_visc=2, scheme=2, check_cfl=3,
                            Re=800.0, Ma=0.3)),
    ("random_rk4_visc07", "random", dict(mx=20, my=12, mz=16, stencil_size=2, stencil_visc=2, scheme=2,
                                         viscexp=0.7, Ma=0.4, Re=300.0, check_cfl=3)),
    ("channel_smr", "channel", CHANNEL),
    ("tgv16_kutta", "tgv", dict(mx=16, my=16, mz=16, stencil_size=2, stencil_visc=2, scheme=1, check_cfl=3)),
    ("tgv20_s3_rk4", "tgv", dict(mx=20, my=20, mz=20, stencil_size=3, stencil_visc=3, scheme=2, check_cfl=3)),
]


def _make(kw, kind, precision):
    cfg = cns.Config(**kw).replace(precision=precision)
    orc = oracle_from(cns.Config(**kw))
    f = initial_fields(kind, orc)
    s = cns.Solver(cfg)
    s.set_state(*f)
    return s


def _compare(precision, splits, tune=False):
    nsteps = sum(splits)
    alone = []
    for name, kind, kw in MEMBERS:
        s = _make(kw, kind, precision)
        s.step(nsteps, outputs=False)
        alone.append((s.get_state(), s.scalars()))
        s.close()
    batch_members = [_make(kw, kind, precision) for _, kind, kw in MEMBERS]
    b = cns.Batch(batch_members)
    groups = b.tune() if tune else b.groups()
    flush_print(f"[case] batch precision={precision} splits={splits} groups={groups}")
    assert len(groups) == 5  # the two RK4 Taylor-Green boxes share a group
    for k in splits:
        b.step(k)
    b.close()
    for (name, _, _), s, (st_a, sc_a) in zip(MEMBERS, batch_members, alone):
        _same(name, precision, s.get_state(), st_a, s.scalars(), sc_a)
        s.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_batch_equals_alone(precision):
    _compare(precision, [1, 3, 3])


def test_batch_tuned_equals_alone():
    _compare(0, [4, 3], tune=True)


def test_members_move_between_batches():
    """A member keeps its step count across batches: A and B take 3 steps in
    one batch, then 4 more in a second batch joined by C, whose count starts
    at 0; each equals its lone run (7, 7 and 4 steps)."""
    sel = [MEMBERS[3], MEMBERS[0], MEMBERS[5]]  # SMR channel (alternating sets), RK4, RK4 s=3
    want = []
    for (name, kind, kw), n in zip(sel, (7, 7, 4)):
        s = _make(kw, kind, 0)
        s.step(n, outputs=False)
        want.append(s.get_state())
        s.close()
    a, b, c = (_make(kw, kind, 0) for _, kind, kw in sel)
    with cns.Batch([a, b]) as b1:
        b1.step(3)
    with cns.Batch([b, c, a]) as b2:
        b2.step(2)
        b2.step(2)
    for (name, _, _), s, w in zip(sel, (a, b, c), want):
        _same(name, 0, s.get_state(), w)
        s.close()


def test_batch_members_are_locked():
    s = _make(MEMBERS[0][2], "tgv", 0)
    b = cns.Batch([s])
    with pytest.raises(cns.CnsError):
        s.step(1, outputs=False)
    with pytest.raises(cns.CnsError):
        cns.Batch([s])
    b.step(1)
    b.close()
    s.step(1, outputs=False)  # free again
    s.close()
