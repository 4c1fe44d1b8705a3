"""Seeded synthetic workloads for the ensemble and serving measurements.

Draws use SplitMix64 with the uniform double (x >> 11) * 2^-53, the same
generator as cns_fill_channel, so every size, parameter and arrival time is
reproducible from its seed alone.
"""
import math

import numpy as np

from .solver import channel, taylor_green

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & _M64

    def next(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & _M64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_in(self, a, b):
        return a + (b - a) * self.uniform()

    def choice(self, seq):
        return seq[min(int(self.uniform() * len(seq)), len(seq) - 1)]

    def int_in(self, a, b):
        """uniform integer in [a, b]"""
        return a + min(int(self.uniform() * (b - a + 1)), b - a)


def ensemble(n_instances=64, seed=4, sizes=(32, 48, 64, 96, 128), order=6, scheme="rk4", precision=1,
             steps=100):
    """Independent Taylor-Green instances: edge length uniform over `sizes`,
    Re uniform in [400, 3200], Mach uniform in [0.1, 0.5]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_instances):
        n = rng.choice(sizes)
        re = rng.uniform_in(400.0, 3200.0)
        ma = rng.uniform_in(0.1, 0.5)
        out.append(dict(config=taylor_green(n=n, order=order, scheme=scheme, Re=re, Ma=ma, precision=precision),
                        kind="tgv", steps=steps))
    return out


def serving(n_requests=1000, seed=5, rate=20.0, order=6, precision=1, tgv_fraction=0.7):
    """Poisson arrival stream of mixed requests (arrival times in seconds):
    Taylor-Green 48^3..160^3, Re 400-3200, Mach 0.1-0.4, 50-400 steps, or an
    isothermal-wall channel of 96x64x64..192x128x96 points (x wall-normal),
    Mach 0.5-1.5, Re 1000-3000, 100-300 steps; all low-storage RK3."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += -math.log(1.0 - rng.uniform()) / rate
        if rng.uniform() < tgv_fraction:
            n = 16 * rng.int_in(3, 10)  # 48 .. 160
            cfg = taylor_green(n=n, order=order, scheme="rk3", Re=rng.uniform_in(400.0, 3200.0),
                               Ma=rng.uniform_in(0.1, 0.4), precision=precision)
            kind, steps = "tgv", rng.int_in(50, 400)
        else:
            f = rng.uniform()
            mx = 16 * int(round(6 + 6 * f))   # 96 .. 192 wall-normal
            my = 16 * int(round(4 + 4 * f))   # 64 .. 128
            mz = 16 * int(round(4 + 2 * f))   # 64 .. 96
            cfg = channel(mx=mx, my=my, mz=mz, order=order, scheme="rk3", Ma=rng.uniform_in(0.5, 1.5),
                          Re=rng.uniform_in(1000.0, 3000.0), precision=precision)
            kind, steps = "channel", rng.int_in(100, 300)
        out.append(dict(id=i, arrival=t, config=cfg, kind=kind, steps=steps, seed=1000 + i))
    return out


def fields_for(solver, kind, seed=1):
    """Initial fields of a request: the Taylor-Green vortex, or the laminar
    channel profile with seeded 10% velocity perturbations."""
    if kind == "tgv":
        return solver.taylor_green_fields()
    return solver.channel_fields(seed=seed, ubulk=1.0, amp=0.1)


def init_on_device(solver, kind, seed=1):
    """The same initial conditions as fields_for, computed on the device."""
    if kind == "tgv":
        solver.init_taylor_green()
    else:
        solver.init_channel(seed=seed, ubulk=1.0, amp=0.1)


def points(config):
    return config.mx * config.my * config.mz


__all__ = ["SplitMix64", "ensemble", "serving", "fields_for", "init_on_device", "points", "np"]
