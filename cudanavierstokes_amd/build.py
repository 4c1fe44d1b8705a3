"""Build the gfx950 shared library libcns.so in-tree with hipcc.

The kernels are compiled once per advective stencil half width (CNS_S=1..4)
and storage type in parallel, the boundary-layer forms (FP64) once per half
width; the host API is its own translation unit.  Packed FP32 vector
instructions are disabled for all device code.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcns.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CNS_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    units = []
    for s in (4, 3, 2, 1):
        for r in ("double", "float"):
            units.append((os.path.join(CSRC, "cns_kernels.hip"), os.path.join(BUILD, f"cns_kernels_s{s}_{r}.o"),
                          [f"-DCNS_S={s}", f"-DCNS_R={r}"]))
        units.append((os.path.join(CSRC, "cns_kernels_bl.hip"), os.path.join(BUILD, f"cns_kernels_bl_s{s}.o"),
                      [f"-DCNS_S={s}"]))
    units.append((os.path.join(CSRC, "cns_api.hip"), os.path.join(BUILD, "cns_api.o"), []))
    cmds = [[HIPCC] + CFLAGS + extra + ["-c", src, "-o", obj] for src, obj, extra in units]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    objs = [obj for _, obj, _ in units]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
