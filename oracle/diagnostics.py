"""CPU restatement of the state diagnostics (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
diagnostics; the product package never imports this module.

  avg_chan_loops / avg_chan   calcAvgChan (src/init.cpp:150-208): per-x plane
                              averages; avg_chan_loops is the reference's own
                              loop nest (small grids), avg_chan its numpy form
  friction_reynolds_loops /   printRes (src/init.cpp:210-240): average friction
  friction_reynolds           Reynolds number at the bottom wall
  flow_stats                  kinetic energy, enstrophy, mean squared
                              dilatation, mean density (no reference
                              counterpart; definitions in its docstring)

Pins (tests/test_diagnostics.py): the loop forms against the numpy forms
on random fields; closed forms (plane-uniform fields have zero
fluctuations; Re_tau of a row set scales as sqrt of the wall velocity
scale; the Taylor-Green kinetic energy 1/8 and enstrophy 3 k'^2/8 with the
stencil's modified wavenumber k', zero dilatation; a 1D wave's mean squared
dilatation k'^2/2).
"""
import math

import numpy as np


def avg_chan_loops(r, u, v, w, e):
    """calcAvgChan as the reference loops it (every term divided by my, then
    mz); fields of shape (mz, my, mx).  Returns the ten profiles rm, um, vm,
    wm, em, rf, uf, vf, wf, ef (lists of mx floats)."""
    mz, my, mx = r.shape
    out = {k: [0.0] * mx for k in ("rm", "um", "vm", "wm", "em", "rf", "uf", "vf", "wf", "ef")}
    for i in range(mx):
        for k in range(mz):
            for j in range(my):
                out["rm"][i] += r[k, j, i] / my / mz
                out["um"][i] += r[k, j, i] * u[k, j, i] / my / mz
                out["vm"][i] += r[k, j, i] * v[k, j, i] / my / mz
                out["wm"][i] += r[k, j, i] * w[k, j, i] / my / mz
                out["em"][i] += e[k, j, i] / my / mz
    for i in range(mx):
        out["um"][i] /= out["rm"][i]
        out["vm"][i] /= out["rm"][i]
        out["wm"][i] /= out["rm"][i]
        for k in range(mz):
            for j in range(my):
                out["rf"][i] += (r[k, j, i] - out["rm"][i]) ** 2 / my / mz
                out["uf"][i] += (u[k, j, i] - out["um"][i]) ** 2 / my / mz
                out["vf"][i] += (v[k, j, i] - out["vm"][i]) ** 2 / my / mz
                out["wf"][i] += (w[k, j, i] - out["wm"][i]) ** 2 / my / mz
                out["ef"][i] += (e[k, j, i] - out["em"][i]) ** 2 / my / mz
    return out


def avg_chan(r, u, v, w, e):
    """numpy form of avg_chan_loops (plane means over axes 0, 1)"""
    mean = lambda a: a.mean(axis=(0, 1))  # noqa: E731
    rm = mean(r)
    um, vm, wm = mean(r * u) / rm, mean(r * v) / rm, mean(r * w) / rm
    em = mean(e)
    return dict(rm=rm, um=um, vm=vm, wm=wm, em=em, rf=mean((r - rm) ** 2), uf=mean((u - um) ** 2),
                vf=mean((v - vm) ** 2), wf=mean((w - wm) ** 2), ef=mean((e - em) ** 2))


def friction_reynolds_loops(r, w, cF, dx, xp0, Re, viscexp=1.0):
    """printRes's loop (src/init.cpp:214-236): cF the advective first
    derivative weights (cF[it] multiplies f[i+it-s] - f[i+s-it])."""
    mz, my, mx = r.shape
    s = len(cF)
    ret = 0.0
    for k in range(mz):
        for j in range(my):
            muw = math.pow(1.0, viscexp) / Re
            ub = [0.0] * (2 * s + 1)
            for i in range(s, 2 * s + 1):
                ub[i] = w[k, j, i - s]
            for i in range(s):
                ub[i] = w[k, j, s - i - 1]
            dudx = 0.0
            for i in range(s):
                dudx += cF[i] * (ub[i] - ub[2 * s - i]) / dx
            dudx *= xp0
            ut = math.sqrt(muw * abs(dudx) / r[k, j, 0])
            ret += ut * r[k, j, 0] / muw
    return ret / my / mz


def friction_reynolds(r, w, cF, dx, xp0, Re, viscexp=1.0):
    s = len(cF)
    muw = 1.0 ** viscexp / Re
    ub = np.concatenate([w[:, :, s - 1::-1], w[:, :, :s + 1]], axis=2)  # (mz, my, 2s+1)
    dudx = np.zeros(r.shape[:2])
    for i in range(s):
        dudx += cF[i] * (ub[:, :, i] - ub[:, :, 2 * s - i]) / dx
    dudx *= xp0
    ut = np.sqrt(muw * np.abs(dudx) / r[:, :, 0])
    return float((ut * r[:, :, 0] / muw).sum() / r.shape[1] / r.shape[0])


def _d1(f, axis, cVF, inv_h, wall=False):
    """viscous-order central first derivative along axis; cVF[it]
    multiplies f[i+it-sv] - f[i+sv-it]; wall: negated mirror ghosts (x)"""
    sv = len(cVF)
    n = f.shape[axis]
    out = np.zeros_like(f)
    for m in range(1, sv + 1):
        idx_p = np.arange(n) + m
        idx_m = np.arange(n) - m
        if wall:
            def take(idx):
                sg = np.where((idx < 0) | (idx >= n), -1.0, 1.0)
                src = np.where(idx < 0, -1 - idx, np.where(idx >= n, 2 * n - 1 - idx, idx))
                shape = [1] * f.ndim
                shape[axis] = n
                return np.take(f, src, axis=axis) * sg.reshape(shape)
            fp, fm = take(idx_p), take(idx_m)
        else:
            fp, fm = np.take(f, idx_p % n, axis=axis), np.take(f, idx_m % n, axis=axis)
        out += cVF[sv - m] * (fm - fp)
    return out * inv_h


def flow_stats(r, u, v, w, g, cVF, vol, wallx=False, nonuni=False):
    """kinetic energy int(rho |u|^2/2)/int(rho), enstrophy
    int(rho |curl u|^2/2)/int(rho), mean (div u)^2, mean rho; integrals with
    weights dxv[i] dy dz, means over the volume vol = Lx Ly Lz"""
    dxi = g["d_dx"] * (g["xp"] if nonuni else np.ones_like(g["xp"]))
    grad = {}
    for a, f in (("u", u), ("v", v), ("w", w)):
        grad[a + "x"] = _d1(f, 2, cVF, 1.0, wall=wallx) * dxi[None, None, :]
        grad[a + "y"] = _d1(f, 1, cVF, g["d_dy"])
        grad[a + "z"] = _d1(f, 0, cVF, g["d_dz"])
    ox = grad["wy"] - grad["vz"]
    oy = grad["uz"] - grad["wx"]
    oz = grad["vx"] - grad["uy"]
    dil = grad["ux"] + grad["vy"] + grad["wz"]
    wgt = (g["dxv"] / g["d_dy"] / g["d_dz"])[None, None, :]
    mass = (r * wgt).sum()
    kin = (0.5 * r * (u * u + v * v + w * w) * wgt).sum() / mass
    ens = (0.5 * r * (ox * ox + oy * oy + oz * oz) * wgt).sum() / mass
    return dict(kinetic=kin, enstrophy=ens, dil2=(dil * dil * wgt).sum() / vol, rho=mass / vol)


def modified_wavenumber(cF, h, k=1.0):
    """Fourier symbol of the central first derivative: D exp(ikx) = i k' exp(ikx),
    k' = (2/h) sum_m a_m sin(m k h) with a_m = -cF[s-m]"""
    s = len(cF)
    return 2.0 / h * sum(-cF[s - m] * math.sin(m * k * h) for m in range(1, s + 1))
