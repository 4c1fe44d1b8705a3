"""The reference's output directory: field files, grid files, an XDMF index,
and checkpoints that resume a run where it stopped.

Formats (fixed by the reference, so that its post-processing and restart
files interchange with this package's):
  <dir>/<c>.<timestep:07d>.bin  c in r, u, v, w, e: the whole grid as
      native-endian float64 in [mz][my][mx] order, variables (rho, u, v, w,
      rhoE)  (saveFileMPI / readFileMPI, src/comm.cpp:205-279)
  <dir>/x.bin, y.bin, z.bin  node coordinates, float64 (src/init.cpp:78-88)
  an XDMF 2.0 file that indexes the field files of several timesteps as a
      temporal collection on a 3DRectMesh (python-utils/writexmf.py:1-75), for
      ParaView / VisIt
A checkpoint adds ctl.<timestep:07d>.json beside the field files: the
configuration and the time-advance control (time, dt, dp/dz, the pending
viscous maximum, the CFL check phase), which the reference's restart drops.
"""
import dataclasses
import json
import os

import numpy as np

NAMES = ("r", "u", "v", "w", "e")


def field_path(directory, name, timestep):
    if not 0 <= int(timestep) <= 9999999:
        raise ValueError("timestep outside 0..9999999")
    return os.path.join(directory, f"{name}.{int(timestep):07d}.bin")


def save_field(directory, name, timestep, a):
    """one field file from a [mz][my][mx] array"""
    np.ascontiguousarray(a, dtype=np.float64).tofile(field_path(directory, name, timestep))


def load_field(directory, name, timestep, shape):
    """one field file as a float64 array of shape (mz, my, mx); the file must
    hold exactly that many values"""
    path = field_path(directory, name, timestep)
    n = int(np.prod(shape))
    size = os.path.getsize(path)
    if size != 8 * n:
        raise ValueError(f"{path} holds {size} bytes, expected {8 * n}")
    return np.fromfile(path, dtype=np.float64).reshape(shape)


def load_fields(directory, timestep, shape):
    """the five fields (rho, u, v, w, rhoE) of a timestep"""
    return [load_field(directory, c, timestep, shape) for c in NAMES]


def load_grid(directory):
    return [np.fromfile(os.path.join(directory, f"{c}.bin"), dtype=np.float64) for c in "xyz"]


def _floats(v):
    return " ".join(f"{float(a):.9e}" for a in v)


def write_xdmf(path, x, y, z, timesteps, times, names=NAMES, precision=8):
    """XDMF 2.0 index of the field files of `timesteps` (simulation times
    `times`), on the rectilinear grid x, y, z; the field files are named
    relative to the XDMF file, which therefore sits in their directory."""
    timesteps = [int(t) for t in timesteps]
    times = [float(t) for t in times]
    if len(times) != len(timesteps):
        raise ValueError("one time per timestep")
    nx, ny, nz = len(x), len(y), len(z)
    dims = f"{nz} {ny} {nx}"
    xml = f'Format="XML" DataType="Float" Precision="{precision}"'
    binary = f'Format="Binary" DataType="Float" Precision="{precision}" Endian="Native"'
    out = ['<?xml version="1.0" ?>',
           '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>',
           '<Xdmf xmlns:xi="http://www.w3.org/2001/XInclude" Version="2.0">',
           '<Domain>',
           f'  <Topology name="TOPO" TopologyType="3DRectMesh" Dimensions="{dims}"/>',
           '  <Geometry name="GEO" GeometryType="VXVYVZ">']
    for v in (x, y, z):
        out += [f'    <DataItem {xml} Dimensions="{len(v)}">', f'      {_floats(v)}', '    </DataItem>']
    out += ['  </Geometry>',
            '  <Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">',
            '    <Time TimeType="List">',
            f'      <DataItem {xml} Dimensions="{len(times)}">',
            f'        {_floats(times)}',
            '      </DataItem>',
            '    </Time>']
    for t in timesteps:
        out += [f'    <Grid Name="T{t:07d}" GridType="Uniform">',
                '      <Topology Reference="/Xdmf/Domain/Topology[1]"/>',
                '      <Geometry Reference="/Xdmf/Domain/Geometry[1]"/>']
        for c in names:
            out += [f'      <Attribute Name="{c}" Center="Node">',
                    f'        <DataItem {binary} Dimensions="{dims}">',
                    f'          {c}.{t:07d}.bin',
                    '        </DataItem>',
                    '      </Attribute>']
        out += ['    </Grid>']
    out += ['  </Grid>', '</Domain>', '</Xdmf>', '']
    with open(path, "w") as f:
        f.write("\n".join(out))


def ctl_path(directory, timestep):
    return os.path.join(directory, f"ctl.{int(timestep):07d}.json")


def save_checkpoint(solver, directory, timestep, extra=None):
    """field files of the current state plus ctl.<timestep>.json (the
    configuration and time-advance control); `extra` is stored alongside"""
    os.makedirs(directory, exist_ok=True)
    solver.write_fields(directory, timestep)
    rec = dict(timestep=int(timestep), config=dataclasses.asdict(solver.config), control=solver.control())
    if extra is not None:
        rec["extra"] = extra
    with open(ctl_path(directory, timestep), "w") as f:
        json.dump(rec, f, indent=1)


def read_checkpoint(directory, timestep):
    with open(ctl_path(directory, timestep)) as f:
        return json.load(f)


def load_checkpoint(solver, directory, timestep):
    """state and control of a checkpoint into `solver` (whose grid must
    match); without a ctl file, the reference's restart (time zero)"""
    solver.read_fields(directory, timestep)
    if os.path.exists(ctl_path(directory, timestep)):
        solver.set_control(read_checkpoint(directory, timestep)["control"])


def solver_from_checkpoint(directory, timestep, **overrides):
    """a new Solver with the checkpoint's configuration, state and control"""
    from .solver import Config, Solver
    rec = read_checkpoint(directory, timestep)
    cfg = Config(**rec["config"]).replace(**overrides) if overrides else Config(**rec["config"])
    s = Solver(cfg)
    try:
        load_checkpoint(s, directory, timestep)
    except Exception:
        s.close()
        raise
    return s


__all__ = ["NAMES", "field_path", "save_field", "load_field", "load_fields", "load_grid", "write_xdmf",
           "save_checkpoint", "read_checkpoint", "load_checkpoint", "solver_from_checkpoint"]
