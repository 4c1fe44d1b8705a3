// main_ref.cpp -- the reference's main program (src/main.cpp:22-79), call for
// call, over include/cns_compat.hpp and libcns.so.  MPI is stubbed for one
// rank, and the reference's host routines that main.cpp and solverWrapper
// call (init.cpp, comm.cpp: splitComm, initGrid, restartWrapper, calcdt,
// printRes, calcAvgChan, writeField) are minimal stand-ins written here,
// since the product replaces only the GPU entry points of src/main.h.
// Test infrastructure: tests/test_gpu_compat.py runs it in a scratch
// directory and compares fields/*.bin with the same run through the Python
// binding.
//
// Case: Taylor-Green vortex 32^3, order 4, low-storage RK3, CFL 0.5,
// nfiles = 2 files of nsteps = 10 steps.
#include <cmath>
#include <cstdio>
#include <ctime>
#include <iostream>
#include <string>

#define myprec double
#define pRow 1
#define pCol 1
#define stencilSize 2
#define stencilVisc 2
#define Lx (2.0 * M_PI)
#define Ly (2.0 * M_PI)
#define Lz (2.0 * M_PI)
#define mx_tot 32
#define my_tot 32
#define mz_tot 32
#define mx (mx_tot)
#define my (my_tot / pRow)
#define mz (mz_tot / pCol)
#define nsteps 10
#define nfiles 2
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
#define idx(i, j, k) ((k) * mx * my + (j) * mx + (i))
const double stretch = 1.8;
const double TwallTop = 1.0;
const double TwallBot = 1.0;

// ---- MPI for one rank ----
typedef int MPI_Comm;
#define MPI_COMM_WORLD 0
inline int MPI_Init(int *, char ***) { return 0; }
inline int MPI_Comm_rank(MPI_Comm, int *r) { return *r = 0; }
inline int MPI_Comm_size(MPI_Comm, int *n) {
    *n = 1;
    return 0;
}
inline int MPI_Barrier(MPI_Comm) { return 0; }
inline int MPI_Finalize() { return 0; }
inline double MPI_Wtime() { return (double)std::clock() / CLOCKS_PER_SEC; }

class Communicator {
   public:
    int rank = 0, nodeRank = 0, jstart = 0, jend = my, kstart = 0, kend = mz;
};

// ---- host globals of src/main.cpp:12-20 ----
double dt, h_dpdz;
double dx, x[mx], xp[mx], xpp[mx], y[my_tot], z[mz_tot];
double r[mx * my * mz];
double u[mx * my * mz];
double v[mx * my * mz];
double w[mx * my * mz];
double e[mx * my * mz];

// ---- stand-ins for the reference's host routines (init.cpp, comm.cpp) ----
void splitComm(Communicator *rk, int myRank) {
    rk->rank = myRank;
    rk->nodeRank = 0;
}
void initGrid(Communicator) {  // uniform periodic grid, cell centres
    dx = Lx / mx;
    for (int i = 0; i < mx; i++) x[i] = Lx * (0.5 + i) / mx, xp[i] = 1.0, xpp[i] = 0.0;
    for (int j = 0; j < my_tot; j++) y[j] = Ly * (0.5 + j) / my_tot;
    for (int k = 0; k < mz_tot; k++) z[k] = Lz * (0.5 + k) / mz_tot;
}
void restartWrapper(Communicator) {  // initCHIT: V0 = T0 = 1, p0 = Rgas
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                const int g = idx(i, j, k);
                u[g] = std::sin(x[i]) * std::cos(y[j]) * std::cos(z[k]);
                v[g] = -std::cos(x[i]) * std::sin(y[j]) * std::cos(z[k]);
                w[g] = 0.0;
                const double p =
                    Rgas + 1.0 / 16.0 * (std::cos(2.0 * x[i]) + std::cos(2.0 * y[j])) * (std::cos(2.0 * z[k]) + 2.0);
                r[g] = p / Rgas;
                e[g] = p / (gam - 1.0) + 0.5 * r[g] * (u[g] * u[g] + v[g] * v[g] + w[g] * w[g]);
            }
}
void calcdt(Communicator) { dt = 0.0; }
void printRes(Communicator) {}
void calcAvgChan(Communicator) {  // mean density per x (the first column of prof.txt)
    if (std::FILE *fp = std::fopen("prof.txt", "w+")) {
        for (int i = 0; i < mx; i++) {
            double rm = 0.0;
            for (int k = 0; k < mz; k++)
                for (int j = 0; j < my; j++) rm += r[idx(i, j, k)] / my / mz;
            std::fprintf(fp, "%lf\t%lf\n", x[i], rm);
        }
        std::fclose(fp);
    }
}
void writeField(int timestep, Communicator) {  // fields/<q>.%07d.bin, [mz][my][mx] float64
    const char names[5] = {'r', 'u', 'v', 'w', 'e'};
    const double *a[5] = {r, u, v, w, e};
    for (int q = 0; q < 5; q++) {
        char path[64];
        std::snprintf(path, sizeof(path), "fields/%c.%07d.bin", names[q], timestep);
        std::FILE *f = std::fopen(path, "wb");
        if (!f) throw std::runtime_error(std::string("cannot open ") + path);
        std::fwrite(a[q], sizeof(double), (size_t)mx * my * mz, f);
        std::fclose(f);
    }
}

#include "cns_compat.hpp"

int main(int argc, char **argv) {
    try {
        // ---- src/main.cpp:22-79 ----
        int myRank, nProcs;
        int ierr;
        ierr = MPI_Init(&argc, &argv);
        ierr = MPI_Comm_rank(MPI_COMM_WORLD, &myRank);
        ierr = MPI_Comm_size(MPI_COMM_WORLD, &nProcs);
        double begin = MPI_Wtime();
        if (nProcs != pRow * pCol) {
            if (myRank == 0) printf("Error! -> nProcs different that pRow*pCol\n");
            ierr = MPI_Barrier(MPI_COMM_WORLD);
            ierr = MPI_Finalize();
            return 0;
        }
        Communicator rk;
        splitComm(&rk, myRank);
        initGrid(rk);
        if (boundaryLayer) calculateSponge(rk);
        restartWrapper(rk);
        calcdt(rk);
        printRes(rk);
        calcAvgChan(rk);
        setDevice(rk.nodeRank);
        setGPUParameters(rk);
        initSolver(rk);
        copyField(0, rk);
        if (restartFile < 0) writeField(0, rk);
        solverWrapper(rk);
        double end = MPI_Wtime();
        std::cout << "The total time is: " << end - begin << "\n";
        double timePerTimeStep = (end - begin) / (nfiles * nsteps);
        std::cout << "The simulation time per time step is: " << timePerTimeStep << "\n";
        std::cout << "Simulation is finished! \n";
        ierr = MPI_Finalize();
        (void)ierr;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
