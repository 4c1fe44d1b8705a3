"""Host-side mirror of the reference solver interface over libcns.so.

Reference entry points (src/main.h:23-42) and their counterparts here:

  setDevice(int)                 -> setDevice(dev)           (config.device)
  setGPUParameters(Communicator) -> setGPUParameters(config) (grid, metrics, coefficients)
  initSolver(Communicator)       -> initSolver()             (device buffers)
  copyField(direction, ...)      -> copyField(0, fields) host->device, copyField(1) device->host
  runSimulation(kin, enst, time) -> runSimulation(nsteps)    (returns kin, enst, time)
  writeField(timestep)           -> writeField(timestep, directory="fields")
  initField(timestep)            -> initField(timestep, directory="fields")
  clearSolver(Communicator)      -> clearSolver()

Solver is the handle-based form of the same interface; several instances can
coexist, unlike the reference's compile-time globals.
"""
import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CnsConfig, CnsControl, CnsError, check, check_batch

SCHEMES = {"rk3": 0, "smr_rk3": 0, "lowstorage": 0, "kutta_rk3": 1, "rk4": 2}


@dataclass
class Config:
    """Run-time counterpart of the reference's case macros (src/globals.h)."""
    mx: int = 64
    my: int = 64
    mz: int = 64
    stencil_size: int = 2
    stencil_visc: int = 2
    periodic_x: int = 1
    nonuniform_x: int = 0
    forcing: int = 0
    sponge: int = 0
    scheme: int = 0
    check_cfl: int = 10
    check_bulk: int = 10
    fix_y_work: int = 0
    precision: int = 0  # 0 FP64 storage, 1 FP32 storage (offsets from a reference state)
    compensated: int = 1  # FP32 storage: compensated (hi + lo) RK sums
    device: int = 0
    Lx: float = 2 * np.pi
    Ly: float = 2 * np.pi
    Lz: float = 2 * np.pi
    Re: float = 1600.0
    Pr: float = 0.71
    Ma: float = 0.1
    gam: float = 1.4
    viscexp: float = 1.0
    cfl: float = 0.5
    stretch: float = 1.8
    twall_bot: float = 1.0
    twall_top: float = 1.0
    dpdz0: float = 0.00372
    boundary_layer: int = 0  # boundaryLayer (FP64, lone solvers, mx a multiple of 16)
    perturbed: int = 0  # perturbed: wall-normal blowing at the wall (with boundary_layer)

    def to_c(self):
        c = CnsConfig()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "scheme" and isinstance(v, str):
                v = SCHEMES[v]
            setattr(c, _lib.C_NAMES.get(f.name, f.name), v)
        return c

    def replace(self, **kw):
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return Config(**d)


def taylor_green(n=64, order=4, scheme="rk3", **kw):
    """Periodic Taylor-Green box of side 2*pi (configs of src/init.cpp:126-148)."""
    s = order // 2
    return Config(mx=n, my=n, mz=n, stencil_size=s, stencil_visc=s, scheme=SCHEMES.get(scheme, scheme), **kw)


def channel(mx=160, my=128, mz=192, order=6, scheme="rk3", **kw):
    """Isothermal-wall channel: x wall-normal (stretched, walls), y spanwise,
    z streamwise with the bulk-velocity controller (globals/channel.h)."""
    s = order // 2
    d = dict(mx=mx, my=my, mz=mz, stencil_size=s, stencil_visc=s, periodic_x=0, nonuniform_x=1,
             forcing=1, scheme=SCHEMES.get(scheme, scheme), Lx=2.0, Ly=4 * np.pi / 3, Lz=4 * np.pi,
             Re=3000.0, Pr=0.72, Ma=1.5, viscexp=0.7, stretch=1.8)
    d.update(kw)
    return Config(**d)


def boundary_layer(mx=240, my=64, mz=2048, order=6, visc_order=4, scheme="rk3", **kw):
    """Spatially developing boundary layer, the reference's default case
    (src/globals.h): x wall-normal (adiabatic wall at x = 0, clustered with
    the half-tanh grid, stretch 5), y spanwise, z streamwise; order 6
    advective / 4 viscous, Lx 20, Ly 7, Lz 500, Re 1500, Pr 0.75, Ma 0.35,
    viscexp 1.5, CFL 0.75, the sponge and the wall-normal blowing on.  The
    sponge strengths and its reference state are set separately
    (inflow.calculate_sponge)."""
    d = dict(mx=mx, my=my, mz=mz, stencil_size=order // 2, stencil_visc=visc_order // 2, periodic_x=0,
             nonuniform_x=1, boundary_layer=1, perturbed=1, sponge=1, forcing=0,
             scheme=SCHEMES.get(scheme, scheme), check_cfl=100, check_bulk=100, Lx=20.0, Ly=7.0, Lz=500.0,
             Re=1500.0, Pr=0.75, Ma=0.35, viscexp=1.5, cfl=0.75, stretch=5.0)
    d.update(kw)
    return Config(**d)


def _f64(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Solver:
    """One flow instance on one GPU (a cns_solver handle)."""

    def __init__(self, config=None, **kw):
        self.config = config if config is not None else Config(**kw)
        self.L = _lib.load()
        self._c = self.config.to_c()
        h = C.c_void_p()
        check(self.L.cns_create(C.byref(self._c), C.byref(h)), None, "cns_create")
        self.h = h
        c = self.config
        self.shape = (c.mz, c.my, c.mx)
        self.n = c.mx * c.my * c.mz

    def close(self):
        if getattr(self, "h", None):
            self.L.cns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        check(rc, self.h, what)

    def grid(self):
        c = self.config
        x, xp, xpp, dxv = (np.empty(c.mx) for _ in range(4))
        y, z = np.empty(c.my), np.empty(c.mz)
        self._chk(self.L.cns_get_grid(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, xp.ctypes.data,
                                      xpp.ctypes.data, dxv.ctypes.data), "cns_get_grid")
        return dict(x=x, y=y, z=z, xp=xp, xpp=xpp, dxv=dxv)

    def taylor_green_fields(self):
        out = [np.empty(self.n) for _ in range(5)]
        self._chk(self.L.cns_fill_taylor_green(self.h, *out), "cns_fill_taylor_green")
        return [a.reshape(self.shape) for a in out]

    def channel_fields(self, seed=1, ubulk=1.0, amp=0.1):
        out = [np.empty(self.n) for _ in range(5)]
        self._chk(self.L.cns_fill_channel(self.h, seed, ubulk, amp, *out), "cns_fill_channel")
        return [a.reshape(self.shape) for a in out]

    def init_taylor_green(self):
        """the Taylor-Green state computed on the device (no host arrays)"""
        self._chk(self.L.cns_init_taylor_green(self.h), "cns_init_taylor_green")

    def init_channel(self, seed=1, ubulk=1.0, amp=0.1):
        """the seeded channel state computed on the device"""
        self._chk(self.L.cns_init_channel(self.h, seed, ubulk, amp), "cns_init_channel")

    def set_state(self, r, u, v, w, e):
        self._chk(self.L.cns_set_state(self.h, *[_f64(a, self.n) for a in (r, u, v, w, e)]), "cns_set_state")

    def get_state(self):
        out = [np.empty(self.n) for _ in range(5)]
        self._chk(self.L.cns_get_state(self.h, *out), "cns_get_state")
        return [a.reshape(self.shape) for a in out]

    def rhs(self):
        out = [np.empty(self.n) for _ in range(5)]
        self._chk(self.L.cns_rhs(self.h, *out), "cns_rhs")
        return [a.reshape(self.shape) for a in out]

    def step(self, nsteps, outputs=True):
        if not outputs:
            self._chk(self.L.cns_step(self.h, nsteps, None, None, None), "cns_step")
            return None
        k, e, t = np.empty(nsteps), np.empty(nsteps), np.empty(nsteps)
        self._chk(self.L.cns_step(self.h, nsteps, k.ctypes.data, e.ctypes.data, t.ctypes.data), "cns_step")
        return k, e, t

    def set_sponge(self, sigma_x, sigma_z, refs):
        c = self.config
        self._chk(self.L.cns_set_sponge(self.h, _f64(sigma_x, c.mx), _f64(sigma_z, c.mz),
                                        *[_f64(a, c.mx * c.mz) for a in refs]), "cns_set_sponge")

    def scalars(self):
        dt, dp, t = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.cns_get_scalars(self.h, C.byref(dt), C.byref(dp), C.byref(t)), "cns_get_scalars")
        return dict(dt=dt.value, dpdz=dp.value, time=t.value)

    def sync(self):
        self._chk(self.L.cns_sync(self.h), "cns_sync")

    def profile_enable(self, on=True):
        self._chk(self.L.cns_profile_enable(self.h, int(on)), "cns_profile_enable")

    def profile_read(self):
        ms, n = C.c_double(), C.c_int()
        self._chk(self.L.cns_profile_read(self.h, C.byref(ms), C.byref(n)), "cns_profile_read")
        return ms.value, n.value

    def time_steps(self, nsteps):
        ms = C.c_double()
        self._chk(self.L.cns_time_steps(self.h, nsteps, C.byref(ms)), "cns_time_steps")
        return ms.value

    def set_pencil(self, pencil):
        """z planes per workgroup march of the stage kernels"""
        self._chk(self.L.cns_set_pencil(self.h, int(pencil)), "cns_set_pencil")

    def tune(self):
        """time a few steps per pencil length, keep the fastest; the state
        is restored bit for bit.  Returns the chosen length."""
        p = C.c_int()
        self._chk(self.L.cns_tune(self.h, C.byref(p)), "cns_tune")
        return p.value

    def device_bytes(self):
        b = C.c_uint64()
        self._chk(self.L.cns_device_bytes(self.h, C.byref(b)), "cns_device_bytes")
        return b.value

    PROFILE_COLUMNS = ("x", "rm", "um", "vm", "wm", "em", "rf", "uf", "vf", "wf", "ef")

    def channel_profiles(self):
        """wall-normal profiles of calcAvgChan (src/init.cpp:150-208): dict of
        the prof.txt columns, each of length mx"""
        out = np.empty(11 * self.config.mx)
        self._chk(self.L.cns_channel_profiles(self.h, out), "cns_channel_profiles")
        return dict(zip(self.PROFILE_COLUMNS, out.reshape(11, self.config.mx)))

    def friction_reynolds(self):
        """average friction Reynolds number of printRes (src/init.cpp:210-240)"""
        r = C.c_double()
        self._chk(self.L.cns_friction_reynolds(self.h, C.byref(r)), "cns_friction_reynolds")
        return r.value

    def flow_stats(self):
        """kinetic energy and enstrophy (density-weighted means), mean squared
        dilatation and mean density of the current state"""
        out = np.empty(4)
        self._chk(self.L.cns_flow_stats(self.h, out), "cns_flow_stats")
        return dict(kinetic=out[0], enstrophy=out[1], dil2=out[2], rho=out[3])

    def step_stats(self, nsteps, every):
        """nsteps steps (as nsteps/every cns_step(every) calls) with the flow
        statistics of the state after each group reduced on the device:
        array (nsteps/every, 4) of kinetic energy, enstrophy, mean squared
        dilatation, mean density"""
        out = np.empty(4 * (nsteps // every))
        self._chk(self.L.cns_step_stats(self.h, int(nsteps), int(every), out), "cns_step_stats")
        return out.reshape(-1, 4)

    def write_fields(self, directory, timestep):
        """<directory>/{r,u,v,w,e}.<timestep:07d>.bin, float64 [mz][my][mx]
        (writeField, src/init.cpp:24-30)"""
        self._chk(self.L.cns_write_fields(self.h, os.fsencode(directory), int(timestep)), "cns_write_fields")

    def read_fields(self, directory, timestep):
        """state from the five field files of a timestep (initField,
        src/init.cpp:13-22); time restarts at zero"""
        self._chk(self.L.cns_read_fields(self.h, os.fsencode(directory), int(timestep)), "cns_read_fields")

    def write_grid(self, directory):
        """<directory>/{x,y,z}.bin node coordinates (src/init.cpp:78-88)"""
        self._chk(self.L.cns_write_grid(self.h, os.fsencode(directory)), "cns_write_grid")

    def control(self):
        """time, dt, dpdz, the viscous maximum pending for the next CFL
        check, the batched-step CFL phase and whether the next check takes
        the viscous limit from the state"""
        c = CnsControl()
        self._chk(self.L.cns_get_control(self.h, C.byref(c)), "cns_get_control")
        return {f: getattr(c, f) for f, _ in CnsControl._fields_ if f != "reserved"}

    def set_control(self, ctl):
        c = CnsControl()
        for f, _ in CnsControl._fields_:
            if f != "reserved":
                setattr(c, f, ctl[f])
        self._chk(self.L.cns_set_control(self.h, C.byref(c)), "cns_set_control")


def mem_info(device=0):
    """(free, total) bytes of device memory (checkGpuMem's query)"""
    f, t = C.c_uint64(), C.c_uint64()
    check(_lib.load().cns_mem_info(int(device), C.byref(f), C.byref(t)), None, "cns_mem_info")
    return f.value, t.value


def release_cache(device=-1):
    """Free the device blocks that closed solvers and batches left cached for
    reuse (of one device, or of all for device < 0)."""
    check(_lib.load().cns_release_cache(int(device)), None, "cns_release_cache")


class Batch:
    """Many solvers advanced together (cns_batch_*): one launch per RK stage
    per kernel-template group, one captured graph per step.  A member's
    result equals its own cns_step result to rounding; members cannot be
    stepped alone while the batch exists."""

    def __init__(self, solvers, tune=False):
        self.L = _lib.load()
        self.members = list(solvers)
        arr = (C.c_void_p * len(self.members))(*[s.h for s in self.members])
        h = C.c_void_p()
        check_batch(self.L.cns_batch_create(arr, len(self.members), C.byref(h)), None, "cns_batch_create")
        self.h = h
        if tune:
            self.tune()

    def _chk(self, rc, what):
        check_batch(rc, self.h, what)

    def tune(self):
        self._chk(self.L.cns_batch_tune(self.h), "cns_batch_tune")
        return self.groups()

    def step(self, nsteps):
        self._chk(self.L.cns_batch_step(self.h, nsteps), "cns_batch_step")

    def launch(self, nsteps):
        """queue nsteps steps and return; wait() before reading members"""
        self._chk(self.L.cns_batch_launch(self.h, nsteps), "cns_batch_launch")

    def wait(self):
        self._chk(self.L.cns_batch_wait(self.h), "cns_batch_wait")

    def time_steps(self, nsteps):
        ms = C.c_double()
        self._chk(self.L.cns_batch_time_steps(self.h, nsteps, C.byref(ms)), "cns_batch_time_steps")
        return ms.value

    def set_shape(self, group, pencil, grid, remap=1):
        """launch shape of one group by hand (instead of tune())"""
        self._chk(self.L.cns_batch_set_shape(self.h, group, pencil, grid, remap), "cns_batch_set_shape")

    def groups(self):
        out = []
        for g in range(self.L.cns_batch_groups(self.h)):
            v = [C.c_int() for _ in range(5)]
            self._chk(self.L.cns_batch_group_info(self.h, g, *[C.byref(x) for x in v]), "cns_batch_group_info")
            out.append(dict(zip(("members", "pencil", "grid", "remap", "work_items"), [x.value for x in v])))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.cns_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Reference-named module functions over one default solver (src/main.h:23-42)
# ---------------------------------------------------------------------------
_default = {"device": 0, "config": None, "solver": None}


def setDevice(device):
    _default["device"] = int(device)


def setGPUParameters(config):
    _default["config"] = config.replace(device=_default["device"])


def initSolver():
    if _default["config"] is None:
        raise CnsError("initSolver: call setGPUParameters first")
    if _default["solver"] is not None:
        _default["solver"].close()
    _default["solver"] = Solver(_default["config"])
    return _default["solver"]


def _solver():
    s = _default["solver"]
    if s is None:
        raise CnsError("initSolver has not been called")
    return s


def copyField(direction, fields_=None):
    """direction 0: host (r,u,v,w,e) -> device; 1: device -> host, returned."""
    s = _solver()
    if direction == 0:
        s.set_state(*fields_)
        return None
    if direction == 1:
        return s.get_state()
    raise ValueError("direction must be 0 or 1")


def runSimulation(nsteps):
    """nsteps steps of the configured scheme; returns (kin, enst, time)."""
    return _solver().step(nsteps)


def writeField(timestep, directory="fields"):
    """the five field files of the current state (src/init.cpp:24-30)"""
    _solver().write_fields(directory, timestep)


def initField(timestep, directory="fields"):
    """state from the field files of a timestep (src/init.cpp:13-22)"""
    _solver().read_fields(directory, timestep)


def clearSolver():
    s = _default["solver"]
    if s is not None:
        s.close()
    _default["solver"] = None
