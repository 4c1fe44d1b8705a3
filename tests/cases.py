"""Initial conditions for the tests, written independently of the product
library (numpy restatement of src/init.cpp:94-148 and of the seeded channel
case), on the grid returned by the oracle."""
import numpy as np

from oracle import oracle as O


def oracle_cfg(**kw):
    return O.make_config(**kw)


def tgv_fields(g, cfg):
    """Taylor-Green vortex of initCHIT (V0 = T0 = R0 = 1, P0 = Rgas)."""
    x, y, z = g["x"], g["y"], g["z"]
    Rg = g["Rgas"]
    fx = 2 * np.pi * x / cfg.Lx
    fy = 2 * np.pi * y / cfg.Ly
    fz = 2 * np.pi * z / cfg.Lz
    Z, Y, X = np.meshgrid(fz, fy, fx, indexing="ij")
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    p = Rg + 1.0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    r = p / Rg
    e = p / (cfg.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    return [r, u, v, w, e]


def splitmix64_uniform(seed, count):
    """SplitMix64 stream, u = (x >> 11) * 2^-53; the state after the i-th
    draw is seed + i*0x9E3779B97F4A7C15 (mod 2^64), so it vectorises."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def channel_fields(g, cfg, seed=1, ubulk=1.0, amp=0.1):
    """Laminar Poiseuille profile along z plus seeded perturbations."""
    mx, my, mz = cfg.mx, cfg.my, cfg.mz
    n = mx * my * mz
    U = splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    x = g["x"]
    xi = x / (cfg.Lx / 2.0) - 1.0
    wm = 1.5 * ubulk * (1.0 - xi * xi)
    i = np.arange(n) % mx
    u = amp * ubulk * (2.0 * U[:, 0] - 1.0)
    v = amp * ubulk * (2.0 * U[:, 1] - 1.0)
    w = wm[i] + amp * ubulk * (2.0 * U[:, 2] - 1.0)
    r = np.ones(n)
    e = g["Rgas"] / (cfg.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    shp = (mz, my, mx)
    return [a.reshape(shp) for a in (r, u, v, w, e)]


def random_smooth_fields(g, cfg, seed=0, amp=0.05):
    """Smooth random periodic perturbation of a uniform state (low modes)."""
    rng = np.random.default_rng(seed)
    x, y, z = g["x"], g["y"], g["z"]
    Z, Y, X = np.meshgrid(2 * np.pi * z / cfg.Lz, 2 * np.pi * y / cfg.Ly, 2 * np.pi * x / cfg.Lx, indexing="ij")

    def field():
        f = np.zeros_like(X)
        for _ in range(4):
            kx, ky, kz = rng.integers(0, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += rng.uniform(-1, 1) * np.cos(kx * X + ph[0]) * np.cos(ky * Y + ph[1]) * np.cos(kz * Z + ph[2])
        return f

    Rg = g["Rgas"]
    r = 1.0 + amp * field()
    u, v, w = field(), field(), field()
    T = 1.0 + amp * field()
    p = r * Rg * T
    e = p / (cfg.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    return [r, u, v, w, e]


def rel_errors(a, b):
    """Relative L2 and Linf per conserved variable of (r, m_x, m_y, m_z, E)
    given two states or RHS sets; the three momentum-like components are
    normalised jointly by the norm of the reference vector."""
    out = {}
    names = ["r", "ru", "rv", "rw", "re"]
    mom = np.sqrt(sum(np.sum(b[q] ** 2) for q in (1, 2, 3)))
    mom_inf = max(np.max(np.abs(b[q])) for q in (1, 2, 3))
    for q, nm in enumerate(names):
        d = np.asarray(a[q]) - np.asarray(b[q])
        if q in (1, 2, 3):
            n2, ninf = mom, mom_inf
        else:
            n2, ninf = np.sqrt(np.sum(b[q] ** 2)), np.max(np.abs(b[q]))
        n2 = n2 if n2 > 0 else 1.0
        ninf = ninf if ninf > 0 else 1.0
        out[nm] = (float(np.sqrt(np.sum(d ** 2)) / n2), float(np.max(np.abs(d)) / ninf))
    return out


def conserved(state):
    r, u, v, w, e = state
    return [r, r * u, r * v, r * w, e]
