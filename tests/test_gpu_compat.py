"""The reference's call sequence from C++ over include/cns_compat.hpp and
libcns.so, bitwise against the same run through the Python binding:
  - tests/compat/main_like.cpp: setDevice, setGPUParameters, initSolver,
    copyField, runSimulation, clearSolver with host arrays;
  - tests/compat/main_ref.cpp: src/main.cpp:22-79 call for call (MPI stubbed
    for one rank), solverWrapper's nfiles loop with its field files and
    solution.txt, checkGpuMem."""
import json
import os
import subprocess

import numpy as np
import pytest

import cudanavierstokes_amd as cns
from tests.helpers import flush_print

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "compat", "main_like")


def _read(path, n):
    a = np.fromfile(path, dtype=np.float64)
    assert a.size == 5 * n
    return [a[q * n:(q + 1) * n].reshape(32, 32, 32) for q in range(5)]


def test_main_like_driver_through_the_shim(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    flush_print("[case] compat driver 32^3 order 4 SMR RK3, 20 steps")
    out = str(tmp_path / "run")
    r = subprocess.run([EXE, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    n = 32 ** 3
    init, final = _read(out + ".init.bin", n), _read(out + ".final.bin", n)
    cfg = cns.taylor_green(n=32, order=4, scheme="rk3", cfl=0.5, check_cfl=10, check_bulk=10, Re=1600.0,
                           Pr=0.71, Ma=0.1, viscexp=1.0)
    with cns.Solver(cfg) as s:
        s.set_state(*init)
        kin, _, time = s.step(20)
        got = s.get_state()
    for a, b in zip(got, final):
        assert np.array_equal(a, b)
    assert info["time"] == time[-1] and info["kin0"] == kin[0]


REF = os.path.join(ROOT, "tests", "compat", "main_ref")


def test_reference_main_through_the_shim(tmp_path):
    if not os.path.exists(REF):
        pytest.fail(f"{REF} is not built (run __graft_entry__.build())")
    flush_print("[case] src/main.cpp-shaped driver: 32^3 order 4 SMR RK3, 2 files x 10 steps")
    (tmp_path / "fields").mkdir()
    r = subprocess.run([REF], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Memory used:" in r.stdout and "Simulation is finished!" in r.stdout
    n = 32 ** 3

    def fields(step):
        return [np.fromfile(tmp_path / "fields" / f"{q}.{step:07d}.bin", dtype=np.float64).reshape(32, 32, 32)
                for q in "ruvwe"]

    init = fields(0)
    cfg = cns.taylor_green(n=32, order=4, scheme="rk3", cfl=0.5, check_cfl=10, check_bulk=10, Re=1600.0,
                           Pr=0.71, Ma=0.1, viscexp=1.0)
    with cns.Solver(cfg) as s:
        s.set_state(*init)
        for f in (1, 2):
            kin, enst, time = s.step(10)
            got = s.get_state()
            for a, b in zip(got, fields(f)):
                assert a.size == n and np.array_equal(a, b), f
    # solution.txt: one line per file (CFL step 0 of each file's 10 steps)
    lines = (tmp_path / "solution.txt").read_text().split()
    assert len(lines) == 2 * 5
    assert float(lines[5 + 1]) == pytest.approx(time[0], abs=1e-6)  # printed with %lf
    assert (tmp_path / "prof.txt").exists()


def test_step_outputs_into_device_arrays():
    """cns_step writes kin/enst/time into device arrays as into host arrays
    (the reference's runSimulation takes device arrays, src/main.h:35).
    Device memory comes from the HIP runtime libcns.so itself uses."""
    import ctypes as C

    from cudanavierstokes_amd import _lib
    lib = _lib.load()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    cfg = cns.taylor_green(n=16, order=4, scheme="rk3", check_cfl=3, check_bulk=2)
    flush_print("[case] cns_step into device arrays, 16^3, 8 steps")
    with cns.Solver(cfg) as a, cns.Solver(cfg) as b:
        a.init_taylor_green()
        b.init_taylor_green()
        kin, enst, time = a.step(8)
        dev = [C.c_void_p() for _ in range(3)]
        for d in dev:
            assert hip.hipMalloc(C.byref(d), 8 * 8) == 0
        try:
            assert lib.cns_step(b.h, 8, dev[0], dev[1], dev[2]) == 0
            b.sync()
            got = [np.empty(8) for _ in range(3)]
            for g, d in zip(got, dev):
                assert hip.hipMemcpy(g.ctypes.data, d, 8 * 8, 2) == 0  # device to host
        finally:
            for d in dev:
                hip.hipFree(d)
        np.testing.assert_array_equal(got[2], time)
        np.testing.assert_array_equal(got[0], kin)
        np.testing.assert_array_equal(np.isnan(got[1]), np.isnan(enst))
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y)
