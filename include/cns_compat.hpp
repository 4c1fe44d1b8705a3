// cns_compat.hpp -- header-only C++ shim that re-exposes the reference's
// solver entry points (src/main.h:23-42) over one default cns_solver, so a
// main.cpp-shaped driver (src/main.cpp:62-69) links against libcns.so
// unchanged.  The caller still provides the reference's compile-time case
// macros (mx, my, mz, stencilSize, ...) and its host globals r, u, v, w, e.
//
// Reference entry point            here
//   setDevice(int)                  setDevice(int)                 (cuda_utils.cu:815)
//   setGPUParameters(Communicator)  setGPUParameters(rk)           (cuda_utils.cu:49)
//   initSolver(Communicator)        initSolver(rk)                 (cuda_utils.cu:415)
//   copyField(int, Communicator)    copyField(direction, rk)       (cuda_utils.cu:299)
//   runSimulation(kin,enst,time,rk) runSimulation(kin,enst,time,rk) (cuda_main.cu:44,109)
//   clearSolver(Communicator)       clearSolver(rk)                (cuda_utils.cu:521)
// Differences: kin/enst/time are host arrays of nsteps entries (the
// reference passes device arrays), failures throw std::runtime_error
// instead of exit(1), and multi-GPU halo exchange is not part of this build.
#ifndef CNS_COMPAT_HPP_
#define CNS_COMPAT_HPP_

#include <stdexcept>
#include <string>

#include "cns.h"

namespace cns_compat {

struct Defaults {
    int device = 0;
    cns_config cfg{};
    bool have_cfg = false;
    cns_solver *solver = nullptr;
};
inline Defaults &state() {
    static Defaults d;
    return d;
}
inline void check(int rc, const char *what) {
    if (rc != CNS_OK)
        throw std::runtime_error(std::string(what) + ": " + cns_last_error(state().solver));
}

}  // namespace cns_compat

inline void setDevice(int device) { cns_compat::state().device = device; }

// Fills the run-time configuration from the reference's case macros.
template <class Communicator>
inline void setGPUParameters(Communicator) {
    cns_config c;
    cns_config_init(&c);
    c.nx = mx_tot;
    c.ny = my_tot;
    c.nz = mz_tot;
    c.stencil_size = stencilSize;
    c.stencil_visc = stencilVisc;
    c.periodic_x = periodicX ? 1 : 0;
    c.nonuniform_x = nonUniformX ? 1 : 0;
    c.forcing_z = forcing ? 1 : 0;
    c.sponge = boundaryLayer ? 1 : 0;
    c.scheme = lowStorage ? CNS_RK3_LOWSTORAGE : CNS_RK3_KUTTA;
    c.check_cfl = checkCFLcondition;
    c.check_bulk = checkBulk;
    c.len_x = Lx;
    c.len_y = Ly;
    c.len_z = Lz;
    c.reynolds = Re;
    c.prandtl = Pr;
    c.mach = Ma;
    c.gamma = gam;
    c.visc_exp = viscexp;
    c.cfl = CFL;
    c.stretch = stretch;
    c.twall_bot = TwallBot;
    c.twall_top = TwallTop;
    c.device = cns_compat::state().device;
    cns_compat::state().cfg = c;
    cns_compat::state().have_cfg = true;
}

template <class Communicator>
inline void initSolver(Communicator) {
    auto &d = cns_compat::state();
    if (!d.have_cfg) throw std::runtime_error("initSolver: call setGPUParameters first");
    if (d.solver) cns_destroy(d.solver);
    d.solver = nullptr;
    cns_compat::check(cns_create(&d.cfg, &d.solver), "initSolver");
}

// direction 0: host globals -> device; 1: device -> host globals
template <class Communicator>
inline void copyField(int direction, Communicator) {
    auto *s = cns_compat::state().solver;
    if (direction == 0)
        cns_compat::check(cns_set_state(s, r, u, v, w, e), "copyField(0)");
    else
        cns_compat::check(cns_get_state(s, r, u, v, w, e), "copyField(1)");
}

template <class Communicator>
inline void runSimulation(double *kin, double *enst, double *time, Communicator) {
    cns_compat::check(cns_step(cns_compat::state().solver, nsteps, kin, enst, time), "runSimulation");
}

template <class Communicator>
inline void clearSolver(Communicator) {
    auto &d = cns_compat::state();
    if (d.solver) cns_destroy(d.solver);
    d.solver = nullptr;
}

#endif  // CNS_COMPAT_HPP_
