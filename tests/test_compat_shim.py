"""include/cns_compat.hpp re-exposes the reference's entry points; it must
compile inside a translation unit that already carries the reference's case
macros (src/globals.h style), which is how a maintainer would drop it in."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASE = r"""
#define myprec double
#define pRow 1
#define pCol 1
#define stencilSize 3
#define stencilVisc 3
#define Lx (2.0)
#define Ly (2.0*M_PI)
#define Lz (4.0*M_PI)
#define mx_tot 160
#define my_tot 128
#define mz_tot 192
#define nsteps 100
#define nfiles 1
#define CFL 0.5f
#define restartFile -1
#define lowStorage (true)
#define boundaryLayer (false)
#define perturbed (false)
#define forcing (true)
#define periodicX (false)
#define nonUniformX (true)
#define checkCFLcondition 10
#define checkBulk 10
#define Re 3000.0
#define Pr 0.72
#define Ma 1.5
#define viscexp 0.7
#define gam 1.4
#define Ec ((gam - 1.f)*Ma*Ma)
#define Rgas (1.f/(gam*Ma*Ma))
#define mx (mx_tot)
#define my (my_tot/pRow)
#define mz (mz_tot/pCol)
#define idx(i,j,k) ({ ( k )*mx*my +( j )*mx + ( i ); })
#include <math.h>
const double stretch = 1.8;
const myprec TwallTop = 1.0;
const myprec TwallBot = 1.0;
double r[mx*my*mz], u[mx*my*mz], v[mx*my*mz], w[mx*my*mz], e[mx*my*mz];
class Communicator { public: int nodeRank, rank, jp, jm, kp, km, jstart, jend, kstart, kend; };
#include "cns_compat.hpp"
int main() {
    Communicator rk{};
    double kin[nsteps], enst[nsteps], time[nsteps];
    setDevice(0);
    setGPUParameters(rk);
    initSolver(rk);
    copyField(0, rk);
    runSimulation(kin, enst, time, rk);
    copyField(1, rk);
    clearSolver(rk);
    return 0;
}
"""


def test_compat_shim_compiles_after_reference_macros(tmp_path):
    src = tmp_path / "main_like.cpp"
    src.write_text(CASE)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "cns.h"\nint main(void){cns_config c; return cns_config_init(&c);}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
