"""Launch shape of the stage kernels: results do not depend on the pencil
length beyond rounding (every point follows the same arithmetic; the
compiler may contract multiply-adds differently in the copies of the plane
loop it generates, so a workgroup's first plane can round differently from
the others -- FP64 relative L2 <= 1e-12, Linf <= 1e-10, FP32 1e-6 / 1e-5),
and cns_tune leaves state, time and step control exactly (bitwise) as it
found them."""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests.helpers import assert_close, flush_print, state_errors

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
    l2, linf = (1e-12, 1e-10) if cfg.precision == 0 else (1e-6, 1e-5)
    ref, sc, _ = _run(cfg, kind)
    for p in (1, 3, 8, cfg.mz):
        got, sg, _ = _run(cfg, kind, pencil=p)
        assert_close(state_errors(got, ref), l2, linf, f"pencil {p}")
        assert abs(sg["time"] - sc["time"]) <= l2 * sc["time"]
    got, sg, chosen = _run(cfg, kind, tune=True)
    assert 1 <= chosen <= cfg.mz
    assert_close(state_errors(got, ref), l2, linf, "after tune")
