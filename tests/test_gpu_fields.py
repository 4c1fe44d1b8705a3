"""Field files on the device path: what cns_write_fields writes is the state
cns_get_state returns, byte for byte, in the reference's format
(src/comm.cpp:205-279); cns_read_fields injects a state written by the
oracle's side and the run matches the oracle; a run resumed from a
checkpoint (field files + time-advance control) matches the uninterrupted
run to rounding (FP64: relative L2 <= 1e-12, Linf <= 1e-10; FP32 storage
1e-6 / 1e-5), alone, in the serving loop and through the command-line
driver."""
import json
import os

import numpy as np
import pytest

import cudanavierstokes_amd as cns
from cudanavierstokes_amd import fields as ff
from cudanavierstokes_amd import run as runmod
from cudanavierstokes_amd.serve import Server
from cudanavierstokes_amd.workloads import init_on_device
from tests.helpers import assert_close, flush_print, initial_fields, oracle_from, state_errors

pytestmark = pytest.mark.gpu


def _tol(cfg):
    return (1e-12, 1e-10) if cfg.precision == 0 else (1e-6, 1e-5)


def _chan(**kw):
    return cns.channel(mx=24, my=16, mz=20, order=6, scheme="rk3", Ma=1.2, Re=2000.0, check_cfl=4, **kw)


@pytest.mark.parametrize("precision", [0, 1])
def test_written_files_hold_the_state(tmp_path, precision):
    cfg = _chan(precision=precision)
    flush_print(f"[case] write fields {cfg}")
    with cns.Solver(cfg) as s:
        s.init_channel(seed=3)
        s.step(3, outputs=False)
        want = s.get_state()
        s.write_fields(str(tmp_path), 3)
        s.write_grid(str(tmp_path))
        g = s.grid()
    for c, w in zip(ff.NAMES, want):
        p = ff.field_path(str(tmp_path), c, 3)
        assert os.path.getsize(p) == 8 * w.size
        assert np.array_equal(np.fromfile(p, dtype=np.float64).reshape(w.shape), w)
    x, y, z = ff.load_grid(str(tmp_path))
    assert np.array_equal(x, g["x"]) and np.array_equal(y, g["y"]) and np.array_equal(z, g["z"])


def test_read_fields_round_trip_and_errors(tmp_path):
    cfg = cns.taylor_green(n=16, order=4, scheme="rk4")
    flush_print(f"[case] read fields {cfg}")
    with cns.Solver(cfg) as s:
        s.init_taylor_green()
        s.step(2, outputs=False)
        want = s.get_state()
        s.write_fields(str(tmp_path), 2)
        with cns.Solver(cfg) as t:
            t.read_fields(str(tmp_path), 2)
            got = t.get_state()
            assert t.scalars()["time"] == 0.0  # the reference's restart
            # u = (rho u) / rho and back rounds once
            assert_close(state_errors(got, want), 1e-15, 1e-14, "read_fields")
            # a truncated file is an error and leaves the state alone
            p = ff.field_path(str(tmp_path), "v", 2)
            with open(p, "r+b") as f:
                f.truncate(8 * 100)
            with pytest.raises(cns.CnsError, match="bytes"):
                t.read_fields(str(tmp_path), 2)
            with pytest.raises(cns.CnsError, match="cannot open"):
                t.read_fields(str(tmp_path), 5)
            assert np.array_equal(t.get_state()[0], got[0])
            with pytest.raises(cns.CnsError, match="cannot open"):
                t.write_fields(str(tmp_path / "no_such_dir"), 1)


def test_injected_channel_matches_oracle(tmp_path):
    """the oracle's seeded channel fields go in through the field files"""
    cfg = _chan()
    flush_print(f"[case] injected channel {cfg}")
    orc = oracle_from(cfg)
    f = initial_fields("channel", orc, seed=1)
    shape = (cfg.mz, cfg.my, cfg.mx)
    for c, a in zip(ff.NAMES, f):
        ff.save_field(str(tmp_path), c, 0, np.asarray(a).reshape(shape))
    orc.set_state(*f)
    with cns.Solver(cfg) as s:
        s.read_fields(str(tmp_path), 0)
        s.step(6, outputs=False)
        got = s.get_state()
    orc.step(6)
    assert_close(state_errors(got, orc.get_state()), 1e-12, 1e-10, "channel from field files vs oracle")


RESUME = [
    ("channel_fp64", _chan(), "channel"),
    ("channel_fp32", _chan(precision=1), "channel"),
    ("tgv_rk4_fp64", cns.taylor_green(n=16, order=4, scheme="rk4", check_cfl=3), "tgv"),
]


@pytest.mark.parametrize("name,cfg,kind", RESUME, ids=[r[0] for r in RESUME])
def test_resume_from_checkpoint(tmp_path, name, cfg, kind):
    flush_print(f"[case] resume {name}")
    with cns.Solver(cfg) as a:
        init_on_device(a, kind, 5)
        a.step(6, outputs=False)
        ff.save_checkpoint(a, str(tmp_path), 6)
        ctl = a.control()
        a.step(5, outputs=False)
        want, want_sc = a.get_state(), a.scalars()
    with ff.solver_from_checkpoint(str(tmp_path), 6) as b:
        got_ctl = b.control()
        assert got_ctl["time"] == ctl["time"] and got_ctl["dpdz"] == ctl["dpdz"]
        assert got_ctl["visc_max"] == ctl["visc_max"] and got_ctl["visc_from_state"] == 0
        b.step(5, outputs=False)
        got, got_sc = b.get_state(), b.scalars()
    l2, linf = _tol(cfg)
    assert_close(state_errors(got, want), l2, linf, f"resumed {name}")
    # dt follows the state: equal to the state's rounding
    assert abs(got_sc["time"] - want_sc["time"]) <= l2 * want_sc["time"]
    if cfg.forcing:
        assert abs(got_sc["dpdz"] - want_sc["dpdz"]) <= max(l2, 1e-9) * abs(want_sc["dpdz"])


def test_served_in_two_parts(tmp_path):
    """a request served for 6 steps, checkpointed, and served again from the
    checkpoint for 5 more equals the lone 11-step run"""
    cfg = _chan()
    flush_print(f"[case] served in parts {cfg}")
    other = dict(id=1, arrival=0.0, config=cns.taylor_green(n=16, order=4, check_cfl=4), kind="tgv", steps=4, seed=2)
    first = [dict(id=0, arrival=0.0, config=cfg, kind="channel", steps=6, seed=7), other]
    done, _ = Server(tick_steps=4, checkpoint_dir=str(tmp_path)).run(first, realtime=False)
    ck = {d["id"]: d for d in done}[0]["checkpoint"]
    assert ck["timestep"] == 6
    rec = ff.read_checkpoint(ck["dir"], 6)
    assert rec["control"]["steps"] == 6 and rec["control"]["visc_from_state"] == 0
    second = [dict(id=0, arrival=0.0, config=cfg, kind="channel", steps=5, restart=ck), dict(other, id=1)]
    done, _ = Server(tick_steps=3, keep_results=True).run(second, realtime=False)
    got = {d["id"]: d for d in done}[0]["state"]
    with cns.Solver(cfg) as s:
        s.init_channel(seed=7)
        s.step(11, outputs=False)
        want = s.get_state()
    assert_close(state_errors(got, want), 1e-12, 1e-10, "served in two parts")


def test_driver_restart_continues_the_run(tmp_path):
    common = ["--case", "tgv", "--n", "16", "--order", "4", "--scheme", "rk3", "--nsteps", "4", "--check-cfl", "2",
              "--check-bulk", "2"]
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    assert runmod.main(common + ["--nfiles", "3", "--out", a]) == 0
    assert runmod.main(common + ["--nfiles", "1", "--out", b]) == 0
    assert runmod.main(common + ["--nfiles", "2", "--restart", "1", "--out", b]) == 0
    shape = (16, 16, 16)
    want = ff.load_fields(os.path.join(a, "fields"), 3, shape)
    got = ff.load_fields(os.path.join(b, "fields"), 3, shape)
    assert_close(state_errors(got, want), 1e-12, 1e-10, "driver restart")
    for d in (a, b):
        fdir = os.path.join(d, "fields")
        names = sorted(os.listdir(fdir))
        for t in range(4):
            for c in ff.NAMES:
                assert f"{c}.{t:07d}.bin" in names
        assert {"x.bin", "y.bin", "z.bin", "fields.xmf"} <= set(names)
        lines = open(os.path.join(d, "solution.txt")).read().split("\n")
        assert [int(l.split()[0]) for l in lines if l] == [1, 3, 5, 7, 9, 11]
        assert len(open(os.path.join(d, "Grid.txt")).read().splitlines()) == 16
        prof = np.loadtxt(os.path.join(d, "prof.txt"))
        assert prof.shape == (16, 11)
    t3 = json.load(open(ff.ctl_path(os.path.join(a, "fields"), 3)))["control"]["time"]
    assert t3 > 0
