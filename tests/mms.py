"""Manufactured solutions: the continuous compressible Navier-Stokes
right-hand side, derived symbolically, for convergence tests of the discrete
operators.  Conservative form in the reference's non-dimensionalisation
(src/globals.h:42-48):
  R = 1/(gam Ma^2), Ec = (gam-1) Ma^2, p = rho R T, mu = T^n/Re,
  lambda = mu/(Pr Ec), rhoE = p/(gam-1) + rho|u|^2/2, h = (rhoE + p)/rho,
  d(rho)/dt   = -d_j(rho u_j)
  d(rho u_i)/dt = -d_j(rho u_i u_j) - d_i p + d_j tau_ij  (+ dpdz for i = z)
  d(rhoE)/dt  = -d_j(rho u_j h) + d_j(u_i tau_ij) + d_j(lambda d_j T)  (+ dpdz w)
  tau_ij = mu (d_j u_i + d_i u_j - 2/3 delta_ij d_k u_k)
"""
import numpy as np
import sympy as sp

X, Y, Z = sp.symbols("x y z", real=True)


def continuous_rhs(rho, u, v, w, T, gam, Ma, Re, Pr, n, dpdz=0):
    R = 1 / (gam * Ma ** 2)
    Ec = (gam - 1) * Ma ** 2
    p = rho * R * T
    E = p / (gam - 1) + rho * (u * u + v * v + w * w) / 2
    h = (E + p) / rho
    mu = T ** n / Re
    lam = mu / (Pr * Ec)
    U = [u, v, w]
    Xs = [X, Y, Z]
    dil = sum(sp.diff(U[i], Xs[i]) for i in range(3))
    tau = [[mu * (sp.diff(U[i], Xs[j]) + sp.diff(U[j], Xs[i]) - (sp.Rational(2, 3) * dil if i == j else 0))
            for j in range(3)] for i in range(3)]
    r0 = -sum(sp.diff(rho * U[j], Xs[j]) for j in range(3))
    rm = []
    for i in range(3):
        e = -sum(sp.diff(rho * U[i] * U[j], Xs[j]) for j in range(3)) - sp.diff(p, Xs[i])
        e += sum(sp.diff(tau[i][j], Xs[j]) for j in range(3))
        if i == 2:
            e += dpdz
        rm.append(e)
    re = -sum(sp.diff(rho * U[j] * h, Xs[j]) for j in range(3))
    re += sum(sp.diff(sum(U[i] * tau[i][j] for i in range(3)), Xs[j]) for j in range(3))
    re += sum(sp.diff(lam * sp.diff(T, Xs[j]), Xs[j]) for j in range(3))
    re += dpdz * w
    return [r0] + rm + [re]


def lambdify_on_grid(exprs, x, y, z):
    Zg, Yg, Xg = np.meshgrid(z, y, x, indexing="ij")
    out = []
    for e in exprs:
        f = sp.lambdify((X, Y, Z), e, "numpy")
        val = f(Xg, Yg, Zg)
        out.append(np.broadcast_to(np.asarray(val, dtype=np.float64), Xg.shape).copy())
    return out


def periodic_fields(L=(2 * np.pi, 2 * np.pi, 2 * np.pi)):
    kx, ky, kz = (2 * sp.pi / sp.Float(l) for l in L)
    rho = 1 + sp.Float(0.08) * sp.sin(kx * X) * sp.cos(ky * Y + sp.Float(0.3)) + sp.Float(0.05) * sp.cos(2 * kz * Z) * sp.sin(ky * Y)
    u = sp.Float(0.7) * sp.sin(kx * X) * sp.cos(ky * Y) * sp.cos(kz * Z) + sp.Float(0.2) * sp.cos(2 * ky * Y + kz * Z)
    v = -sp.Float(0.5) * sp.cos(kx * X) * sp.sin(ky * Y) * sp.cos(kz * Z) + sp.Float(0.3) * sp.sin(kx * X + kz * Z)
    w = sp.Float(0.3) * sp.sin(kx * X + ky * Y) * sp.cos(2 * kz * Z) + sp.Float(0.1)
    T = 1 + sp.Float(0.1) * sp.cos(kx * X - kz * Z) * sp.sin(ky * Y)
    return rho, u, v, w, T


def wall_fields(Lx, Ly, Lz, Tw=1.0, R=1.0):
    """Fields consistent with the isothermal no-slip ghost rules at x = 0, Lx:
    velocity odd, pressure even, T - Tw odd, dv/dy + dw/dz = 0 (so the
    dilatation is even, matching its mirrored ghosts)."""
    sx = sp.sin(sp.pi * X / Lx)
    cx = sp.cos(sp.pi * X / Lx)
    ky, kz = 2 * sp.pi / Ly, 2 * sp.pi / Lz
    psi = sp.sin(ky * Y) * sp.sin(kz * Z)
    u = sx * (sp.Float(0.4) + sp.Float(0.2) * sp.cos(ky * Y) * sp.sin(kz * Z))
    v = sp.Float(0.3) * sx * sp.diff(psi, Z)
    w = -sp.Float(0.3) * sx * sp.diff(psi, Y) + sp.Float(0.5) * sx
    p = R * (1 + sp.Float(0.05) * cx * sp.cos(ky * Y))
    T = Tw + sp.Float(0.1) * sx * (1 + sp.Float(0.5) * sp.cos(kz * Z))
    rho = p / (R * T)
    return rho, u, v, w, T


def state_arrays(fields, x, y, z, gam, Ma):
    rho, u, v, w, T = fields
    R = 1 / (gam * Ma ** 2)
    E = rho * R * T / (gam - 1) + rho * (u * u + v * v + w * w) / 2
    r, uu, vv, ww, ee = lambdify_on_grid([rho, u, v, w, E], x, y, z)
    return [r, uu, vv, ww, ee]
