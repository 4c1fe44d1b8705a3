"""ctypes wrapper of the CPU FP64 oracle (oracle/cns_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OrcConfig(C.Structure):
    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("s", C.c_int), ("sv", C.c_int),
        ("periodic_x", C.c_int), ("nonuniform_x", C.c_int),
        ("forcing", C.c_int), ("sponge", C.c_int), ("scheme", C.c_int),
        ("check_cfl", C.c_int), ("check_bulk", C.c_int), ("fix_y_work", C.c_int),
        ("boundary_layer", C.c_int), ("perturbed", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("Re", C.c_double), ("Pr", C.c_double), ("Ma", C.c_double),
        ("gam", C.c_double), ("viscexp", C.c_double), ("cfl", C.c_double),
        ("stretch", C.c_double), ("twall_bot", C.c_double), ("twall_top", C.c_double),
        ("dpdz0", C.c_double),
    ]


SCHEMES = {"rk3": 0, "smr_rk3": 0, "kutta_rk3": 1, "rk4": 2}

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_state.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_get_state.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_rhs.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_get_field.argtypes = [vp, C.c_int, dp]
        L.orc_get_grid.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, vp]
        L.orc_set_sponge.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp]
        for f in ("orc_get_dt", "orc_get_dpdz", "orc_get_time", "orc_calc_dt"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [vp]
        L.orc_set_dt.argtypes = [vp, C.c_double]
        L.orc_set_time.argtypes = [vp, C.c_double]
        L.orc_perturbation.restype = C.c_int
        L.orc_perturbation.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.orc_bulk.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_vol_integral.restype = C.c_double
        L.orc_vol_integral.argtypes = [vp, dp, C.c_int]
        L.orc_step.argtypes = [vp, C.c_int, vp, vp, vp]
        L.orc_rk_scalar.restype = C.c_double
        L.orc_rk_scalar.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_coeffs.argtypes = [C.c_int, dp, dp]
        L.orc_pencil_d1.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_pencil_d2.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_pencil_flux.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, vp, dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_config_size.restype = C.c_int
        assert L.orc_config_size() == C.sizeof(OrcConfig)
        _lib = L
    return _lib


def make_config(**kw):
    """OrcConfig from keyword arguments (same names as the struct)."""
    d = dict(mx=16, my=16, mz=16, s=2, sv=2, periodic_x=1, nonuniform_x=0, forcing=0,
             sponge=0, scheme=0, check_cfl=10, check_bulk=10, fix_y_work=0, boundary_layer=0, perturbed=0,
             Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi, Re=1600.0, Pr=0.71, Ma=0.1,
             gam=1.4, viscexp=1.0, cfl=0.5, stretch=1.8, twall_bot=1.0, twall_top=1.0,
             dpdz0=0.00372)
    if "scheme" in kw and isinstance(kw["scheme"], str):
        kw["scheme"] = SCHEMES[kw["scheme"]]
    d.update(kw)
    c = OrcConfig()
    for k, v in d.items():
        setattr(c, k, v)
    return c


FIELD_IDS = {"h": 0, "t": 1, "p": 2, "m": 3, "l": 4, "dudx": 5, "dvdx": 6, "dwdx": 7,
             "dudy": 8, "dvdy": 9, "dwdy": 10, "dudz": 11, "dvdz": 12, "dwdz": 13, "dil": 14}


class Oracle:
    """One oracle solver instance; arrays are float64 in the reference
    layout idx = i + j*mx + k*mx*my, returned with shape (mz, my, mx)."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else make_config(**kw)
        self.L = lib()
        self.h = self.L.orc_create(C.byref(self.cfg))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        c = self.cfg
        self.shape = (c.mz, c.my, c.mx)
        self.n = c.mx * c.my * c.mz

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))

    def set_state(self, r, u, v, w, e):
        self.L.orc_set_state(self.h, *[self._arr(a) for a in (r, u, v, w, e)])

    def get_state(self):
        out = [np.empty(self.n) for _ in range(5)]
        self.L.orc_get_state(self.h, *out)
        return [a.reshape(self.shape) for a in out]

    def rhs(self):
        out = [np.empty(self.n) for _ in range(5)]
        self.L.orc_rhs(self.h, *out)
        return [a.reshape(self.shape) for a in out]

    def field(self, name):
        out = np.empty(self.n)
        self.L.orc_get_field(self.h, FIELD_IDS[name], out)
        return out.reshape(self.shape)

    def grid(self):
        c = self.cfg
        x, xp, xpp, dxv = (np.empty(c.mx) for _ in range(4))
        y, z = np.empty(c.my), np.empty(c.mz)
        sc = np.empty(9)
        vsx = np.empty(c.mx * (2 * c.sv + 1))
        self.L.orc_get_grid(self.h, x, xp, xpp, y, z, dxv, sc, vsx.ctypes.data)
        keys = ["dx", "d_dx", "d_dy", "d_dz", "d_d2x", "d_d2y", "d_d2z", "Rgas", "Ec"]
        g = dict(x=x, xp=xp, xpp=xpp, y=y, z=z, dxv=dxv, vsx=vsx.reshape(2 * c.sv + 1, c.mx))
        g.update({k: float(v) for k, v in zip(keys, sc)})
        return g

    def set_sponge(self, sx, sz, refs):
        self.L.orc_set_sponge(self.h, self._arr(sx), self._arr(sz), *[self._arr(a) for a in refs])

    def step(self, nsteps):
        par1, par2, t = np.empty(nsteps), np.empty(nsteps), np.empty(nsteps)
        self.L.orc_step(self.h, nsteps, par1.ctypes.data, par2.ctypes.data, t.ctypes.data)
        return par1, par2, t

    @property
    def dt(self):
        return self.L.orc_get_dt(self.h)

    @property
    def dpdz(self):
        return self.L.orc_get_dpdz(self.h)

    @property
    def time(self):
        return self.L.orc_get_time(self.h)

    @time.setter
    def time(self, t):
        self.L.orc_set_time(self.h, float(t))

    def perturbation(self, j, k, t):
        """Blowing value of u's bottom ghosts at (j, k) and time t, or None
        outside the band."""
        v = C.c_double()
        return v.value if self.L.orc_perturbation(self.h, int(j), int(k), float(t), C.byref(v)) else None

    def calc_dt(self):
        return self.L.orc_calc_dt(self.h)

    def bulk(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_bulk(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def vol_integral(self, f, average=False):
        return self.L.orc_vol_integral(self.h, self._arr(f), int(average))


def coeffs(s):
    cF, cS = np.zeros(4), np.zeros(5)
    lib().orc_coeffs(s, cF, cS)
    return cF[:s].copy(), cS[:s + 1].copy()


def pencil_d1(f, s, d):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    lib().orc_pencil_d1(len(f), s, d, f, out)
    return out


def pencil_d2(f, s, d2):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    lib().orc_pencil_d2(len(f), s, d2, f, out)
    return out


def pencil_flux(f, g, h, s, d):
    f = np.ascontiguousarray(f, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty_like(f)
    hp = None
    if h is not None:
        h = np.ascontiguousarray(h, dtype=np.float64)
        hp = h.ctypes.data
    lib().orc_pencil_flux(len(f), s, d, f, g, hp, out)
    return out


def rk_scalar(scheme, lam, dt):
    if isinstance(scheme, str):
        scheme = SCHEMES[scheme]
    return lib().orc_rk_scalar(scheme, lam, dt)


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()
