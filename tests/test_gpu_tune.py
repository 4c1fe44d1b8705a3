"""Launch shape of the stage kernels: results do not depend on the pencil
length (bitwise), and cns_tune leaves state, time and step control exactly
as it found them."""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests.helpers import flush_print

pytestmark = pytest.mark.gpu

CASES = [
    ("tgv_fp64_rk4", cns.taylor_green(n=24, order=4, scheme="rk4", check_cfl=3), "tgv"),
    ("channel_fp32_rk3", cns.channel(mx=24, my=16, mz=40, order=6, scheme="rk3", check_cfl=4, precision=1), "channel"),
    ("channel_fp64_kutta", cns.channel(mx=20, my=16, mz=24, order=4, scheme="kutta_rk3", check_cfl=2), "channel"),
]


def _run(cfg, kind, pencil=None, tune=False, steps=5):
    with cns.Solver(cfg) as s:
        if kind == "tgv":
            s.init_taylor_green()
        else:
            s.init_channel(seed=3)
        s.step(2, outputs=False)
        before = (s.get_state(), s.control())
        if pencil is not None:
            s.set_pencil(pencil)
        chosen = s.tune() if tune else None
        if tune:
            after = (s.get_state(), s.control())
            for a, b in zip(before[0], after[0]):
                assert np.array_equal(a, b), "tune changed the state"
            assert before[1] == after[1], "tune changed the step control"
        s.step(steps, outputs=False)
        return s.get_state(), s.scalars(), chosen


@pytest.mark.parametrize("name,cfg,kind", CASES, ids=[c[0] for c in CASES])
def test_pencil_length_does_not_change_results(name, cfg, kind):
    flush_print(f"[case] pencil {name}")
    ref, sc, _ = _run(cfg, kind)
    for p in (1, 3, 8, cfg.mz):
        got, sg, _ = _run(cfg, kind, pencil=p)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), f"pencil {p}"
        assert sg == sc
    got, sg, chosen = _run(cfg, kind, tune=True)
    assert 1 <= chosen <= cfg.mz
    for a, b in zip(ref, got):
        assert np.array_equal(a, b), "after tune"
    assert sg == sc
