"""cns_rhs_noalloc, the entry point that allocates nothing (no counterpart in
the reference; for callers that cannot allocate device memory): one
right-hand side of a state the caller holds on the device, into a device
output the caller holds, with the dilatation field in the library's
file-scope scratch array.  Checked against cns_rhs on the same state (the
same kernels, so bitwise equal), on two calls in a row, the second on a
smaller grid (the scratch keeps the first call's contents), and a grid above
the scratch maximum is an error."""
import ctypes as C

import numpy as np
import pytest

import cudanavierstokes_amd as cns
from cudanavierstokes_amd import _lib
from tests.helpers import flush_print, initial_fields, oracle_from

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _lib_noalloc():
    return _lib.load()


def _compare(cfg, kind):
    flush_print(f"[case] cns_rhs_noalloc {kind} {cfg.mx}x{cfg.my}x{cfg.mz} s={cfg.stencil_size}")
    L, hip = _lib_noalloc(), _hip()
    f = initial_fields(kind, oracle_from(cfg))
    with cns.Solver(cfg) as s:
        s.set_state(*f)
        want = s.rhs()
        r, u, v, w, e = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in f)
        n = r.size
        q = np.concatenate([r, r * u, r * v, r * w, e])  # the device layout: five conserved fields
        got = np.empty(5 * n)
        dq, dr = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(dq), 5 * n * 8) == 0
        assert hip.hipMalloc(C.byref(dr), 5 * n * 8) == 0
        try:
            assert hip.hipMemcpy(dq, q.ctypes.data, 5 * n * 8, H2D) == 0
            rc = L.cns_rhs_noalloc(s.h, dq, dr)
            assert rc == 0, L.cns_last_error(s.h)
            s.sync()
            assert hip.hipMemcpy(got.ctypes.data, dr, 5 * n * 8, D2H) == 0
        finally:
            hip.hipFree(dq)
            hip.hipFree(dr)
    for k in range(5):
        assert np.array_equal(got[k * n:(k + 1) * n].reshape(want[k].shape), want[k]), k


def test_noalloc_rhs_equals_rhs_twice_then_smaller():
    _compare(cns.Config(mx=48, my=40, mz=36, stencil_size=3, stencil_visc=3, periodic_x=0, nonuniform_x=1,
                        forcing=1, Lx=2.0, Ly=4 * np.pi / 3, Lz=4 * np.pi, Re=3000.0, Pr=0.72, Ma=1.5,
                        viscexp=0.7), "channel")
    _compare(cns.Config(mx=20, my=16, mz=12, stencil_size=4, stencil_visc=4), "tgv")


def test_noalloc_rejects_grids_above_the_scratch_maximum():
    L = _lib_noalloc()
    cfg = cns.Config(mx=160, my=128, mz=128, stencil_size=2, stencil_visc=2)  # 2.6 M points, 21 MB of FP64
    with cns.Solver(cfg) as s:
        rc = L.cns_rhs_noalloc(s.h, C.c_void_p(16), C.c_void_p(16))
        assert rc != 0
        assert b"maximum" in L.cns_last_error(s.h)
