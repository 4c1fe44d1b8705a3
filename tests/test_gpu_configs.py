"""The five benchmark workloads at their own sizes, against the FP64 oracle.

  C1  Taylor-Green 64^3, order 4, low-storage RK3, CFL 0.5, 100 steps, FP64
  C2  Taylor-Green 256^3, order 8, classic RK4, 2 steps, FP64 and FP32
      storage, each with the launch shape the benchmark uses (cns_tune)
  C3  isothermal-wall channel 160 (wall-normal x, stretched) x 128 x 192,
      order 6, low-storage RK3 with the bulk-velocity controller, 5 steps,
      FP32 storage (and FP64)
  C4  the 64-instance seeded Taylor-Green ensemble (sizes {32..128}^3,
      order 6, RK4, FP32 storage, 100 steps) as one tuned batch: every member
      against its lone run, four sampled members against the oracle
  C5  the first 20 requests of the seeded serving stream at their real
      sizes and step counts through the serving loop: every request against
      its lone run, three (one of them a channel) against the oracle

Tolerances (relative per conserved variable, momenta normalised jointly):
FP64 against the oracle L2 <= 1e-12, Linf <= 1e-10; FP32 storage against
the FP64 oracle L2 <= 1e-5, Linf <= 1e-4; batched or served against the
lone run 1e-12 / 1e-10 (FP64) and 1e-6 / 1e-5 (FP32).  Time step, time and
dp/dz against the oracle: relative 1e-12 (FP64), 1e-6 (FP32).
"""
import numpy as np
import pytest

import cudanavierstokes_amd as cns
from cudanavierstokes_amd.serve import Server
from cudanavierstokes_amd.workloads import ensemble, init_on_device, serving
from tests.helpers import assert_close, flush_print, initial_fields, oracle_from, state_errors

pytestmark = pytest.mark.gpu

TOL64 = (1e-12, 1e-10)
TOL32 = (1e-5, 1e-4)
ALONE = {0: (1e-12, 1e-10), 1: (1e-6, 1e-5)}


def _oracle_run(cfg, kind, nsteps, seed=1):
    orc = oracle_from(cfg.replace(precision=0))
    f = initial_fields(kind, orc, seed=seed)
    orc.set_state(*f)
    p1, p2, t = orc.step(nsteps)
    out = dict(fields=f, state=orc.get_state(), time=t, dt=orc.dt, dpdz=orc.dpdz, par1=p1)
    orc.close()
    return out


def _gpu_run(cfg, f, nsteps, tune=True):
    s = cns.Solver(cfg)
    s.set_state(*f)
    pencil = s.tune() if tune else None
    kin, enst, t = s.step(nsteps)
    out = dict(state=s.get_state(), time=t, kin=kin, sc=s.scalars(), pencil=pencil)
    s.close()
    return out


def _check(name, cfg, g, o):
    l2, linf = TOL64 if cfg.precision == 0 else TOL32
    errs = state_errors(g["state"], o["state"])
    flush_print(f"{name} pencil={g['pencil']} errors={errs}")
    assert_close(errs, l2, linf, name)
    rs = 1e-12 if cfg.precision == 0 else 1e-6
    np.testing.assert_allclose(g["time"], o["time"], rtol=rs, atol=0)
    assert g["sc"]["dt"] == pytest.approx(o["dt"], rel=rs)
    if cfg.forcing:
        assert g["sc"]["dpdz"] == pytest.approx(o["dpdz"], rel=max(rs, 1e-10))
    # bulk statistic at the bulk-check steps
    m = ~np.isnan(o["par1"])
    np.testing.assert_allclose(g["kin"][m], o["par1"][m], rtol=max(rs, 1e-12))


@pytest.mark.timeout(600)
def test_c1_tgv64_o4_rk3_100_steps():
    cfg = cns.taylor_green(n=64, order=4, scheme="rk3", cfl=0.5, Re=1600.0, Ma=0.1)
    flush_print(f"[case] C1 {cfg}")
    o = _oracle_run(cfg, "tgv", 100)
    _check("C1 fp64", cfg, _gpu_run(cfg, o["fields"], 100), o)


@pytest.fixture(scope="module")
def c2_oracle():
    cfg = cns.taylor_green(n=256, order=8, scheme="rk4", cfl=0.5, check_cfl=10, check_bulk=10, Re=1600.0,
                           Ma=0.1)
    flush_print(f"[case] C2 oracle {cfg}")
    return cfg, _oracle_run(cfg, "tgv", 2)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_c2_tgv256_o8_rk4(c2_oracle, precision):
    base, o = c2_oracle
    cfg = base.replace(precision=precision, compensated=1)
    flush_print(f"[case] C2 precision={precision}")
    _check(f"C2 {'fp32' if precision else 'fp64'}", cfg, _gpu_run(cfg, o["fields"], 2), o)


@pytest.fixture(scope="module")
def c3_oracle():
    cfg = cns.channel(mx=160, my=128, mz=192, order=6, scheme="rk3", check_cfl=10, check_bulk=10)
    flush_print(f"[case] C3 oracle {cfg}")
    return cfg, _oracle_run(cfg, "channel", 5, seed=1)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", [1, 0], ids=["fp32", "fp64"])
def test_c3_channel_160x128x192_o6_rk3(c3_oracle, precision):
    base, o = c3_oracle
    cfg = base.replace(precision=precision)
    flush_print(f"[case] C3 precision={precision}")
    _check(f"C3 {'fp32' if precision else 'fp64'}", cfg, _gpu_run(cfg, o["fields"], 5), o)


def _sample_members(inst):
    """First member of each of the sizes 32, 48, 64 and the second of size 32
    (the oracle needs seconds per member up to 64^3 x 100 RK4 steps)."""
    picks, seen = [], {}
    for k, i in enumerate(inst):
        n = i["config"].mx
        seen[n] = seen.get(n, 0) + 1
        if (n in (48, 64) and seen[n] == 1) or (n == 32 and seen[n] <= 2):
            picks.append(k)
    return picks


@pytest.mark.timeout(900)
def test_c4_ensemble_64_instances():
    steps = 100
    inst = ensemble(n_instances=64, seed=4, steps=steps)
    flush_print(f"[case] C4 64 instances, {steps} steps")
    alone = []
    for i in inst:
        s = cns.Solver(i["config"])
        s.set_state(*s.taylor_green_fields())
        s.step(steps, outputs=False)
        alone.append((s.get_state(), s.scalars()))
        s.close()
    sols = []
    for i in inst:
        s = cns.Solver(i["config"])
        s.set_state(*s.taylor_green_fields())
        sols.append(s)
    b = cns.Batch(sols)
    groups = b.tune()
    flush_print(f"C4 groups {groups}")
    b.step(steps)
    b.close()
    got = [(s.get_state(), s.scalars()) for s in sols]
    for s in sols:
        s.close()
    l2, linf = ALONE[1]
    for k, ((gs, gsc), (ws, wsc)) in enumerate(zip(got, alone)):
        assert_close(state_errors(gs, ws), l2, linf, f"C4 member {k} batch vs alone")
        assert gsc["time"] == pytest.approx(wsc["time"], rel=1e-6)
    for k in _sample_members(inst):
        cfg = inst[k]["config"]
        o = _oracle_run(cfg, "tgv", steps)
        errs = state_errors(got[k][0], o["state"])
        flush_print(f"C4 member {k} ({cfg.mx}^3, Re {cfg.Re:.0f}, Ma {cfg.Ma:.3f}) vs oracle {errs}")
        assert_close(errs, *TOL32, f"C4 member {k} vs oracle")
        assert got[k][1]["time"] == pytest.approx(o["time"][-1], rel=1e-6)


def _oracle_picks(reqs):
    """The two cheapest Taylor-Green requests and the cheapest channel one."""
    work = lambda r: r["config"].mx * r["config"].my * r["config"].mz * r["steps"]
    tg = sorted((r for r in reqs if r["kind"] == "tgv"), key=work)[:2]
    ch = sorted((r for r in reqs if r["kind"] == "channel"), key=work)[:1]
    return [r["id"] for r in tg + ch]


@pytest.mark.timeout(900)
def test_c5_served_stream_real_sizes():
    reqs = serving(n_requests=20, seed=5, rate=20.0)
    flush_print(f"[case] C5 first 20 requests: {[(r['kind'], r['config'].mx, r['steps']) for r in reqs]}")
    done, summary = Server(budget_bytes=200e9, tick_steps=10, keep_results=True).run(reqs, realtime=False)
    flush_print(f"C5 summary {summary}")
    assert summary["requests"] == len(reqs)
    by_id = {d["id"]: d for d in done}
    for r in reqs:
        s = cns.Solver(r["config"])
        init_on_device(s, r["kind"], r["seed"])
        s.step(r["steps"], outputs=False)
        want = s.get_state()
        s.close()
        assert_close(state_errors(by_id[r["id"]]["state"], want), *ALONE[1], f"C5 request {r['id']} vs alone")
    for rid in _oracle_picks(reqs):
        r = reqs[rid]
        o = _oracle_run(r["config"], r["kind"], r["steps"], seed=r["seed"])
        errs = state_errors(by_id[rid]["state"], o["state"])
        flush_print(f"C5 request {rid} ({r['kind']} {r['config'].mx}x{r['config'].my}x{r['config'].mz}, "
                    f"{r['steps']} steps) vs oracle {errs}")
        assert_close(errs, *TOL32, f"C5 request {rid} vs oracle")
