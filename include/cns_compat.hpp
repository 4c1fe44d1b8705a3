// cns_compat.hpp -- header-only C++ shim that re-exposes the reference's GPU
// entry points (src/main.h:23-42) over one default cns_solver, so the
// reference's main program (src/main.cpp:22-79) and its host code (init.cpp,
// comm.cpp: initGrid, restartWrapper, writeField, calcAvgChan, ...) link
// against libcns.so in place of the reference's .cu files.  The caller
// provides, as the reference does, the compile-time case macros of
// src/globals.h (mx, my, mz, stencilSize, nsteps, nfiles, restartFile, ...),
// the host globals r, u, v, w, e and x, z (filled by its initGrid), and the
// host functions writeField(int, Communicator) and calcAvgChan(Communicator)
// that solverWrapper calls.
//
// Reference entry point                        here
//   setDevice(int)                              (src/cuda_utils.cu:815)
//   setGPUParameters(Communicator)              (src/cuda_utils.cu:49)
//   initSolver(Communicator)                    (src/cuda_utils.cu:415)
//   copyField(int, Communicator)                (src/cuda_utils.cu:299)
//   runSimulation(kin, enst, time, rk)          (src/cuda_main.cu:44)
//   runSimulationLowStorage(kin, enst, time, rk) (src/cuda_main.cu:109)
//   solverWrapper(Communicator)                 (src/cuda_main.cu:267-324)
//   clearSolver(Communicator)                   (src/cuda_utils.cu:521)
//   calculateSponge(Communicator)               (src/sponge.cu:83-240)
//   checkGpuMem(Communicator)                   (src/cuda_utils.cu:780-813)
// kin/enst/time may be host or device arrays of nsteps entries.  Failures
// throw std::runtime_error instead of exit(1).  One GPU: the reference's MPI
// halo exchange is not part of this build (pRow = pCol = 1).
//
// calculateSponge reads blasius1D/{x,r,u,w,e}Prof.bin as the reference does
// and passes the sponge to the solver that initSolver creates.  The
// reference computes the sponge but never copies it to the device
// (copySpongeToDevice at src/sponge.cu:41-44 is empty), so its boundary-layer
// runs apply no sponge; this build applies the sponge it computes.
#ifndef CNS_COMPAT_HPP_
#define CNS_COMPAT_HPP_

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "cns.h"

namespace cns_compat {

struct Defaults {
    int device = 0;
    cns_config cfg{};
    bool have_cfg = false;
    cns_solver *solver = nullptr;
    // sponge of calculateSponge, set on the solver by initSolver
    bool have_sponge = false;
    std::vector<double> sx, sz, ref[5];
};
inline Defaults &state() {
    static Defaults d;
    return d;
}
inline void check(int rc, const char *what) {
    if (rc != CNS_OK)
        throw std::runtime_error(std::string(what) + ": " + cns_last_error(state().solver));
}
inline std::vector<double> read_profile(const char *path) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("calculateSponge: cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> a(bytes > 0 ? (size_t)bytes / sizeof(double) : 0);
    const size_t got = std::fread(a.data(), sizeof(double), a.size(), f);
    std::fclose(f);
    if (got != a.size() || a.empty()) throw std::runtime_error(std::string("calculateSponge: reading error in ") + path);
    return a;
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
    c.boundary_layer = boundaryLayer ? 1 : 0;
#ifdef perturbed
    c.wall_blowing = perturbed ? 1 : 0;
#endif
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
    if (d.have_sponge)
        cns_compat::check(cns_set_sponge(d.solver, d.sx.data(), d.sz.data(), d.ref[0].data(), d.ref[1].data(),
                                         d.ref[2].data(), d.ref[3].data(), d.ref[4].data()),
                          "initSolver (sponge)");
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

// nsteps steps; the scheme (low-storage or Kutta RK3) is the one
// setGPUParameters took from lowStorage, under either name
template <class Communicator>
inline void runSimulation(double *kin, double *enst, double *time, Communicator) {
    cns_compat::check(cns_step(cns_compat::state().solver, nsteps, kin, enst, time), "runSimulation");
}
template <class Communicator>
inline void runSimulationLowStorage(double *kin, double *enst, double *time, Communicator rk) {
    runSimulation(kin, enst, time, rk);
}

template <class Communicator>
inline void clearSolver(Communicator) {
    auto &d = cns_compat::state();
    if (d.solver) cns_destroy(d.solver);
    d.solver = nullptr;
}

// Memory report of src/cuda_utils.cu:780-813 (one GPU: no halo buffers).
template <class Communicator>
inline void checkGpuMem(Communicator rk) {
    uint64_t free_b = 0, total_b = 0;
    cns_compat::check(cns_mem_info(rk.nodeRank, &free_b, &total_b), "checkGpuMem");
    const double gib = 1024.0 * 1024.0 * 1024.0;
    const double free_m = (double)free_b / gib, total_m = (double)total_b / gib, used_m = total_m - free_m;
    double overhead = 9 * sizeof(void *) + 20 * 3 * sizeof(unsigned);
    overhead += (7 + mx * (3 + 2 * stencilSize + 2 * stencilVisc + 2) + 2 * stencilSize + 2 * stencilVisc) * sizeof(double);
    const double field = (double)mx * my * mz * sizeof(double);
    const double fields = (used_m * gib - overhead) / field, totfields = (total_m * gib - overhead) / field;
    if (rk.rank == 0) {
        std::printf("Memory usage details: \n");
        std::printf("Memory used: %lf GB\n", used_m);
        std::printf("Memory free: %lf GB\n", free_m);
        std::printf("Number of fields allocatable: %d\n", (int)totfields);
        std::printf("Number of fields allocated:   %d\n", (int)fields);
        std::printf("\n");
    }
}

// Sponge and reference state of a boundary-layer run (src/sponge.cu:83-240):
// the profile of blasius1D/*.bin interpolated onto the grid x, z (host
// globals of initGrid), the zones of src/sponge.h; with restartFile < 0 the
// reference state, constant along y, becomes the initial condition.  Writes
// inProf.txt and inRef.txt as the reference does.
template <class Communicator>
inline void calculateSponge(Communicator rk) {
    extern double x[], z[];  // the grid of initGrid (src/main.cpp:14)
    auto &d = cns_compat::state();
    const std::vector<double> xin = cns_compat::read_profile("blasius1D/xProf.bin");
    const std::vector<double> rin = cns_compat::read_profile("blasius1D/rProf.bin");
    const std::vector<double> uin = cns_compat::read_profile("blasius1D/uProf.bin");
    const std::vector<double> win = cns_compat::read_profile("blasius1D/wProf.bin");
    const std::vector<double> ein = cns_compat::read_profile("blasius1D/eProf.bin");
    const int n = (int)xin.size();
    if ((int)rin.size() != n || (int)uin.size() != n || (int)win.size() != n || (int)ein.size() != n)
        throw std::runtime_error("calculateSponge: blasius1D profiles differ in length");
    d.sx.assign(mx, 0.0);
    d.sz.assign(mz, 0.0);
    for (auto &a : d.ref) a.assign((size_t)mx * mz, 0.0);
    std::vector<double> zl(mz);
    for (int k = 0; k < mz; k++) zl[k] = z[k + rk.kstart];
    cns_compat::check(cns_sponge_from_profile(n, xin.data(), rin.data(), uin.data(), win.data(), mx, x, mz, zl.data(),
                                              Lx, Lz, Re, gam, Ma, nullptr, d.sx.data(), d.sz.data(),
                                              d.ref[0].data(), d.ref[1].data(), d.ref[2].data(), d.ref[3].data(),
                                              d.ref[4].data()),
                      "calculateSponge");
    d.have_sponge = true;
    if (std::FILE *fw = std::fopen("inProf.txt", "w+")) {
        for (int i = 0; i < n; i++) std::fprintf(fw, "%le %le %le %le %le\n", xin[i], rin[i], uin[i], win[i], ein[i]);
        std::fclose(fw);
    }
    if (restartFile < 0) {
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) {
                    const size_t g2 = i + (size_t)k * mx, g = i + (size_t)j * mx + (size_t)k * mx * my;
                    r[g] = d.ref[0][g2];
                    u[g] = d.ref[1][g2] / d.ref[0][g2];
                    v[g] = 0.0;
                    w[g] = d.ref[3][g2] / d.ref[0][g2];
                    e[g] = d.ref[4][g2];
                }
    }
    if (std::FILE *fw = std::fopen("inRef.txt", "w+")) {
        for (int i = 0; i < mx; i++) std::fprintf(fw, "%le %le %le %le %le\n", x[i], r[i], u[i], w[i], e[i]);
        std::fclose(fw);
    }
}

// The run loop of src/cuda_main.cu:267-324: nfiles times nsteps steps, each
// followed by copyField(1), writeField(file) and calcAvgChan (the caller's
// host functions), a progress line, and the CFL-step lines of solution.txt
// (file*(t+1), time, par1, par2, dt); then clearSolver.
template <class Communicator>
inline void solverWrapper(Communicator rk) {
    std::vector<double> hpar1(nsteps), hpar2(nsteps), htime(nsteps);
    checkGpuMem(rk);
    const int start = restartFile < 0 ? 0 : restartFile;
    std::FILE *fp = rk.rank == 0 ? std::fopen("solution.txt", "w+") : nullptr;
    for (int file = start + 1; file < nfiles + start + 1; file++) {
        if (lowStorage)
            runSimulationLowStorage(hpar1.data(), hpar2.data(), htime.data(), rk);
        else
            runSimulation(hpar1.data(), hpar2.data(), htime.data(), rk);
        copyField(1, rk);
        writeField(file, rk);
        calcAvgChan(rk);
        if (rk.rank == 0) {
            std::printf("file number: %d  \t step: %d  \t time: %lf  \t kin: %le  \t energy: %le\n", file,
                        file * nsteps, htime[nsteps - 1], hpar1[nsteps - 1], hpar2[nsteps - 1]);
            if (fp)
                for (int t = 0; t < nsteps - 1; t += checkCFLcondition)
                    std::fprintf(fp, "%d %lf %lf %lf %lf\n", file * (t + 1), htime[t], hpar1[t], hpar2[t],
                                 htime[t + 1] - htime[t]);
        }
    }
    if (fp) std::fclose(fp);
    clearSolver(rk);
}

#endif  // CNS_COMPAT_HPP_
