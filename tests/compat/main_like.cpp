// main_like.cpp -- a driver in the shape of the reference's main program
// (src/main.cpp:22-75): case macros as in src/globals.h, host arrays
// r, u, v, w, e filled as initCHIT does (src/init.cpp:126-148), then the
// reference's call sequence through include/cns_compat.hpp on libcns.so.
// Test infrastructure: tests/test_gpu_compat.py runs it and compares with the
// same run through the Python binding.
//
//   main_like <out_prefix>   writes <out_prefix>.init.bin and <out_prefix>.final.bin
//                            (r, u, v, w, e, each mx*my*mz float64) and prints
//                            kin[nsteps-1] and time[nsteps-1]
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#define myprec double
#define stencilSize 2
#define stencilVisc 2
#define Lx (2.0 * M_PI)
#define Ly (2.0 * M_PI)
#define Lz (2.0 * M_PI)
#define mx_tot 32
#define my_tot 32
#define mz_tot 32
#define mx mx_tot
#define my my_tot
#define mz mz_tot
#define nsteps 20
#define nfiles 1
#define restartFile -1
#define CFL 0.5
#define lowStorage (true)
#define boundaryLayer (false)
#define forcing (false)
#define periodicX (true)
#define nonUniformX (false)
#define checkCFLcondition 10
#define checkBulk 10
#define Re 1600.0
#define Pr 0.71
#define Ma 0.1
#define viscexp 1.0
#define gam 1.4
#define Rgas (1.0 / (gam * Ma * Ma))
const double stretch = 1.8;
const double TwallTop = 1.0;
const double TwallBot = 1.0;

static double r[mx * my * mz], u[mx * my * mz], v[mx * my * mz], w[mx * my * mz], e[mx * my * mz];

class Communicator {
   public:
    int nodeRank = 0, rank = 0;
};

#include "cns_compat.hpp"

static void dump(const std::string &path) {
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + path);
    for (const double *a : {r, u, v, w, e}) std::fwrite(a, sizeof(double), mx * my * mz, f);
    std::fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: main_like <out_prefix>\n");
        return 2;
    }
    const std::string out = argv[1];
    Communicator rk;
    // initCHIT: Taylor-Green vortex, V0 = T0 = 1, p0 = Rgas
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                const double x = Lx * (0.5 + i) / mx, y = Ly * (0.5 + j) / my, z = Lz * (0.5 + k) / mz;
                const int g = i + j * mx + k * mx * my;
                u[g] = std::sin(x) * std::cos(y) * std::cos(z);
                v[g] = -std::cos(x) * std::sin(y) * std::cos(z);
                w[g] = 0.0;
                const double p = Rgas + 1.0 / 16.0 * (std::cos(2.0 * x) + std::cos(2.0 * y)) * (std::cos(2.0 * z) + 2.0);
                r[g] = p / Rgas;
                e[g] = p / (gam - 1.0) + 0.5 * r[g] * (u[g] * u[g] + v[g] * v[g] + w[g] * w[g]);
            }
    dump(out + ".init.bin");
    try {
        std::vector<double> kin(nsteps), enst(nsteps), time(nsteps);
        setDevice(rk.nodeRank);
        setGPUParameters(rk);
        initSolver(rk);
        copyField(0, rk);
        runSimulation(kin.data(), enst.data(), time.data(), rk);
        copyField(1, rk);
        clearSolver(rk);
        dump(out + ".final.bin");
        std::printf("{\"kin0\": %.17g, \"time\": %.17g}\n", kin[0], time[nsteps - 1]);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
