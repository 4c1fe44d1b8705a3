"""Pairs an oracle instance with a product solver on identical settings."""
import numpy as np

from oracle import oracle as O
from tests import cases

# product Config field -> oracle field
_MAP = {"stencil_size": "s", "stencil_visc": "sv"}
_COMMON = ["mx", "my", "mz", "periodic_x", "nonuniform_x", "forcing", "sponge", "scheme", "check_cfl",
           "check_bulk", "fix_y_work", "Lx", "Ly", "Lz", "Re", "Pr", "Ma", "gam", "viscexp", "cfl", "stretch",
           "twall_bot", "twall_top", "dpdz0", "boundary_layer", "perturbed"]


def oracle_from(config):
    kw = {k: getattr(config, k) for k in _COMMON}
    kw["s"] = config.stencil_size
    kw["sv"] = config.stencil_visc
    return O.Oracle(**kw)


def initial_fields(kind, orc, seed=1):
    g = orc.grid()
    if kind == "tgv":
        return cases.tgv_fields(g, orc.cfg)
    if kind == "channel":
        return cases.channel_fields(g, orc.cfg, seed=seed)
    if kind == "random":
        return cases.random_smooth_fields(g, orc.cfg, seed=seed)
    raise ValueError(kind)


def flush_print(*a):
    print(*a, flush=True)


def assert_close(errs, l2, linf, what):
    bad = {k: v for k, v in errs.items() if not (v[0] <= l2 and v[1] <= linf)}
    assert not bad, f"{what}: relative (L2, Linf) above ({l2:g}, {linf:g}): {bad}; all: {errs}"


def state_errors(gpu_state, orc_state):
    return cases.rel_errors(cases.conserved(gpu_state), cases.conserved(orc_state))


def max_abs(a):
    return float(np.max(np.abs(a)))
