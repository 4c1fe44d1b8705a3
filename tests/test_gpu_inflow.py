"""calculateSponge on the device path: the sponge strengths and Blasius
reference state of cudanavierstokes_amd.inflow, set with cns_set_sponge and
used as the initial state, give the oracle's right-hand side (relative L2
<= 1e-12, Linf <= 1e-10) and steps."""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from cudanavierstokes_amd import inflow as fl
from tests import cases
from tests.helpers import assert_close, flush_print, oracle_from, state_errors

pytestmark = pytest.mark.gpu


def test_sponge_with_similarity_reference():
    cfg = cns.Config(mx=24, my=8, mz=40, stencil_size=2, stencil_visc=2, periodic_x=0, nonuniform_x=1,
                     sponge=1, Lx=4.0, Ly=2.0, Lz=60.0, Re=300.0, Ma=0.3, Pr=0.72, viscexp=0.7, stretch=1.8,
                     twall_bot=1.0, twall_top=1.0, scheme=0, check_cfl=2)
    flush_print(f"[case] sponge inflow {cfg}")
    with cns.Solver(cfg) as s:
        out = fl.calculate_sponge(s, set_state=True)
        assert out["sigma_x"].max() > 0 and out["sigma_z"].max() > 0
        state0 = s.get_state()
        rg = s.rhs()
        s.step(4, outputs=False)
        sg = s.get_state()
    orc = oracle_from(cfg)
    orc.set_sponge(out["sigma_x"], out["sigma_z"], out["refs"])
    orc.set_state(*state0)
    ro = orc.rhs()
    assert_close(cases.rel_errors(rg, ro), 1e-12, 1e-10, "sponge rhs")
    orc.step(4)
    assert_close(state_errors(sg, orc.get_state()), 1e-12, 1e-10, "sponge steps")
