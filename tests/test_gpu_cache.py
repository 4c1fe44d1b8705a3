"""Device blocks of closed solvers stay cached for the next solver of the same
size (cns_destroy, cns_create) and are freed by cns_release_cache.  A solver
built on reused blocks computes exactly what a solver on fresh blocks does
(whose parity with the oracle the other GPU tests hold)."""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests.helpers import flush_print, initial_fields, oracle_from

pytestmark = pytest.mark.gpu


def _run(cfg, f):
    s = cns.Solver(cfg)
    s.set_state(*f)
    r = s.rhs()
    s.step(3)
    st = s.get_state()
    return s, r, st


def test_closed_solver_blocks_are_reused_then_released():
    cfg = cns.taylor_green(n=96, order=4, scheme="rk3")  # blocks well above the 2-4 MiB allocation granules
    f = initial_fields("tgv", oracle_from(cfg))
    cns.release_cache()
    f0 = cns.mem_info()[0]
    a, ra, sta = _run(cfg, f)  # fresh blocks
    nbytes = a.device_bytes()
    a.close()
    f1 = cns.mem_info()[0]
    flush_print(f"[case] cache: solver {nbytes} B, free before {f0}, after close {f1}")
    assert f0 - f1 >= 0.9 * nbytes, "the closed solver's blocks stay cached"
    b, rb, stb = _run(cfg, f)  # the same blocks again
    f2 = cns.mem_info()[0]
    assert abs(f2 - f1) < 0.1 * nbytes, "a solver of the same size takes the cached blocks"
    for x, y in zip(ra + sta, rb + stb):
        assert np.array_equal(x, y), "reused blocks give the bits of fresh ones"
    b.close()
    cns.release_cache()
    f3 = cns.mem_info()[0]
    assert f3 - f2 >= 0.9 * nbytes, "cns_release_cache frees the cached blocks"
