"""Boundary-layer inflow data: the sponge of calculateSponge
(src/sponge.cu:83-313, constants of src/sponge.h:4-17) with the
compressible self-similar (Blasius) profile of python-utils/selfSimilarSol.py
as its reference state.

  sponge_strengths     sigma_x (top, x near Lx) and sigma_z (inlet z near 0,
                       outlet z near Lz): strength * (distance into the zone
                       / zone length) ** exponent
  similarity_profile   the laminar compressible boundary layer in similarity
                       form, solved as a two-point boundary-value problem;
                       returned on y / delta_99 like blasius1D/*.bin
  natural_spline       cubic spline with zero end curvature (the
                       spline/splint pair of src/sponge.cu:258-313)
  reference_fields     the sponge reference state: the profile stretched with
                       the boundary-layer growth sqrt(1 + z/Re) (src/sponge.cu:158-168)
  calculate_sponge     both, set on a solver (cns_set_sponge); optionally the
                       reference state as initial condition (src/sponge.cu:186-196)

Axes as in the reference: x wall-normal, z streamwise, y spanwise.
Deviations: the reference's spline routine indexes its 0-based arrays from 1
(reading one element past each profile); natural_spline interpolates the
profile as given.  selfSimilarSol.py fixes the viscosity exponent to 3/2
whatever its argument; here exp_mu is honoured (default 3/2).
"""
import math

import numpy as np

# src/sponge.h:4-17: (strength, length, exponent)
SPONGE_TOP = (1.0, 1.0, 2.0)
SPONGE_INLET = (0.5, 20.0, 2.0)
SPONGE_OUTLET = (0.5, 20.0, 2.0)


def sponge_strengths(x, z, Lx, Lz, top=SPONGE_TOP, inlet=SPONGE_INLET, outlet=SPONGE_OUTLET):
    """sigma_x[mx], sigma_z[mz] (src/sponge.cu:112-129); the outlet zone
    wins where the inlet and outlet zones overlap"""
    x, z = np.asarray(x, dtype=float), np.asarray(z, dtype=float)
    sx = np.zeros_like(x)
    s, L, e = top
    if L > 0:
        m = x >= Lx - L
        sx[m] = s * ((x[m] - (Lx - L)) / L) ** e
    sz = np.zeros_like(z)
    s, L, e = inlet
    if L > 0:
        m = z <= L
        sz[m] = s * ((L - z[m]) / L) ** e
    s, L, e = outlet
    if L > 0:
        m = z >= Lz - L
        sz[m] = s * ((z[m] - (Lz - L)) / L) ** e
    return sx, sz


def similarity_profile(gam=1.4, Ma=0.3, Pr=0.72, exp_mu=1.5, n=600, eta_max=10.0, n_out=1000, dy_tail=0.15):
    """Laminar compressible boundary layer over an adiabatic wall in
    similarity variables (python-utils/selfSimilarSol.py:21-60): with the
    stream function F, U = F', T the temperature, C = T**(exp_mu - 1),

        F' = U,  (C U')' = -F (C U'),
        (C T'/Pr)' = -F (C T'/Pr) - Ec C U'**2,   y' = sqrt(2) T,

    F(0) = U(0) = T'(0) = y(0) = 0, U -> 1 and T -> 1 far away, Ec =
    (gam - 1) Ma**2.  Returns y / delta_99 and the density 1/T, the
    wall-normal velocity V, the streamwise velocity U and the internal energy
    T Rgas / (gam - 1), Rgas = 1 / (gam Ma**2), on n_out points: the solution
    points, then constant values every dy_tail (the layout of blasius1D/*.bin,
    python-utils/selfSimilarSol.py:62-98)."""
    from scipy.integrate import solve_bvp

    Ec = (gam - 1.0) * Ma * Ma
    Rgas = 1.0 / (gam * Ma * Ma)

    def rhs(_, f):
        C = f[3] ** (exp_mu - 1.0)
        return np.vstack((f[1], f[2] / C, -f[0] * f[2] / C, f[4] * Pr / C,
                          -f[0] * f[4] * Pr / C - Ec * f[2] ** 2 / C, math.sqrt(2.0) * f[3]))

    def bc(a, b):
        return np.array([a[0], a[1], a[4], a[5], b[1] - 1.0, b[3] - 1.0])

    eta = np.linspace(0.0, eta_max, n)
    f0 = np.zeros((6, n))
    f0[1] = np.sqrt(eta / eta[-1])
    f0[3] = 1.0
    f0[5] = eta
    sol = solve_bvp(rhs, bc, eta, f0, tol=1e-10, max_nodes=200000)
    if not sol.success:
        raise RuntimeError(f"similarity solution did not converge: {sol.message}")
    F, U, G, T, H, y = sol.sol(eta)
    delta = y[np.where(U > 0.99)[0][0]]
    r = 1.0 / T
    V = (U * y / 2.0 - F / (r * math.sqrt(2.0))) / delta
    out = {k: np.zeros(n_out) for k in ("y", "r", "u", "w", "e")}
    m = min(n, n_out)
    out["y"][:m] = (y / delta)[:m]
    out["r"][:m] = r[:m]
    out["w"][:m] = U[:m]
    out["u"][:m] = V[:m]
    out["e"][:m] = T[:m] * Rgas / (gam - 1.0)
    for i in range(m, n_out):
        out["y"][i] = out["y"][i - 1] + dy_tail
    for k in ("r", "u", "w", "e"):
        out[k][m:] = out[k][m - 1]
    out.update(delta=float(delta), wall_shear=float(G[0]), wall_temperature=float(T[0]))
    return out


def natural_spline(xk, yk):
    """Interpolating cubic spline with zero second derivative at both ends
    (src/sponge.cu:258-313): returns a vectorised evaluator; outside the
    knots it continues the end interval's cubic."""
    xk = np.asarray(xk, dtype=float)
    yk = np.asarray(yk, dtype=float)
    n = xk.size
    h = np.diff(xk)
    if np.any(h <= 0):
        raise ValueError("knots must increase strictly")
    # tridiagonal system for the interior second derivatives
    M = np.zeros(n)
    if n > 2:
        a = h[:-1]
        b = 2.0 * (h[:-1] + h[1:])
        c = h[1:]
        d = 6.0 * (np.diff(yk[1:]) / h[1:] - np.diff(yk[:-1]) / h[:-1])
        m = n - 2
        cp, dp = np.zeros(m), np.zeros(m)
        cp[0], dp[0] = c[0] / b[0], d[0] / b[0]
        for i in range(1, m):
            den = b[i] - a[i] * cp[i - 1]
            cp[i] = c[i] / den if i < m - 1 else 0.0
            dp[i] = (d[i] - a[i] * dp[i - 1]) / den
        sol = np.zeros(m)
        sol[-1] = dp[-1]
        for i in range(m - 2, -1, -1):
            sol[i] = dp[i] - cp[i] * sol[i + 1]
        M[1:-1] = sol

    def ev(x):
        x = np.asarray(x, dtype=float)
        k = np.clip(np.searchsorted(xk, x, side="right") - 1, 0, n - 2)
        hh = xk[k + 1] - xk[k]
        A = (xk[k + 1] - x) / hh
        B = (x - xk[k]) / hh
        return A * yk[k] + B * yk[k + 1] + ((A ** 3 - A) * M[k] + (B ** 3 - B) * M[k + 1]) * hh * hh / 6.0

    return ev


def reference_fields(x, z, prof, Re, gam, Ma):
    """Sponge reference state on [mz][mx] (src/sponge.cu:158-168): at
    streamwise station z the profile is read at x / s, s = sqrt(1 + z/Re);
    the wall-normal velocity is V / (s Re), the streamwise one U; pressure
    Rgas, so rho E = Rgas/(gam - 1) + rho (u**2 + w**2)/2.  Returns the
    conserved fields (rho, rho u, rho v, rho w, rho E)."""
    x, z = np.asarray(x, dtype=float), np.asarray(z, dtype=float)
    Rgas = 1.0 / (gam * Ma * Ma)
    sp = {k: natural_spline(prof["y"], prof[k]) for k in ("r", "u", "w")}
    scale = np.sqrt(1.0 + z / Re)[:, None]
    xs = x[None, :] / scale
    r = sp["r"](xs)
    u = sp["u"](xs) / (scale * Re)
    w = sp["w"](xs)
    e = Rgas / (gam - 1.0) + r * 0.5 * (u * u + w * w)
    return [r, r * u, np.zeros_like(r), r * w, e]


def write_profile_files(directory, prof):
    """The profile as the reference's blasius1D/{x,r,u,w,e}Prof.bin (raw
    float64; x holds y / delta_99), the input calculateSponge reads
    (src/sponge.cu:131-156; written by python-utils/selfSimilarSol.py)."""
    import os
    os.makedirs(directory, exist_ok=True)
    for name, key in (("x", "y"), ("r", "r"), ("u", "u"), ("w", "w"), ("e", "e")):
        np.ascontiguousarray(prof[key], dtype=np.float64).tofile(os.path.join(directory, f"{name}Prof.bin"))


def calculate_sponge(solver, prof=None, set_state=False, **zones):
    """calculateSponge for a solver: strengths from the zone constants
    (keywords top, inlet, outlet override src/sponge.h), reference state from
    the similarity profile (computed if not given); with set_state the
    reference state, constant along y, becomes the initial condition.
    Returns the arrays it set."""
    c = solver.config
    g = solver.grid()
    if prof is None:
        prof = similarity_profile(gam=c.gam, Ma=c.Ma, Pr=c.Pr)
    sx, sz = sponge_strengths(g["x"], g["z"], c.Lx, c.Lz, **zones)
    refs = reference_fields(g["x"], g["z"], prof, c.Re, c.gam, c.Ma)
    solver.set_sponge(sx, sz, refs)
    if set_state:
        r, ru, _, rw, e = (np.broadcast_to(a[:, None, :], (c.mz, c.my, c.mx)) for a in refs)
        solver.set_state(r, ru / r, np.zeros_like(r), rw / r, e)
    return dict(sigma_x=sx, sigma_z=sz, refs=refs, profile=prof)


__all__ = ["SPONGE_TOP", "SPONGE_INLET", "SPONGE_OUTLET", "sponge_strengths", "similarity_profile", "natural_spline",
           "reference_fields", "write_profile_files", "calculate_sponge"]
