// cns_batch.hpp -- kernels that advance many independent flow instances in
// one launch per RK stage (ensembles and the serving loop).
//
// Every instance keeps its own state, buffers and device scalars; a launch
// covers the work items (x-y tile x z chunk) of all instances of a group that
// share one kernel template (order, boundary kind, storage precision).  The
// per-step control of cns_step (CFL check every check_cfl steps, viscous
// limit from the state on the first check, forcing controller, time) is kept
// per instance on the device, so one captured sequence of launches is valid
// for every step of every instance.
#pragma once
#include "cns_aux.hpp"
#include "cns_rhs.hpp"

namespace cns {

// Per-instance step control, on the device next to Scalars.
struct StepCtl {
    int istep;       // steps taken since the instance joined
    int vfs;         // next CFL check takes the viscous limit from the state
    int check_cfl;   // CFL / forcing check period
    int forcing;     // dp/dz controller on
    double cfl;      // CFL number
    double vol;      // Lx Ly Lz
};

// One instance in one stage launch.  A[p] is the stage's operand set for
// step parity p (they differ only for the low-storage RK3, whose state
// alternates between two sets).
template <typename R>
struct Job {
    Params P;
    StagePtrs<R> A[2];
    R *dil;
    Scalars *sc;
    StepCtl *ctl;
    double *part;  // NPART * RED_BLOCKS partials of this instance
    int first;     // first work item of this instance in the launch
    int nwork;     // work items (tiles x chunks) of this instance
};

// Job arrays are read through the constant address space: nothing in a
// stage launch writes them (nor the step counters, which only the control
// kernels between stage launches advance), so every read is a scalar load
// and the values live in scalar registers -- as the tiles' buffer resources
// require (a field pointer the compiler took for divergent would need a
// loop over its distinct values around every buffer access).
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

// stage arrays of a job for step parity par
template <typename R>
__device__ __forceinline__ StagePtrs<R> stage_arrays(cptr<Job<R>> J, int par) {
    StagePtrs<R> g;
    g.qin = J->A[par].qin;
    g.dil = J->A[par].dil;
    g.s1 = J->A[par].s1;
    g.ra = J->A[par].ra;
    g.rb = J->A[par].rb;
    g.s2 = J->A[par].s2;
    g.qout = J->A[par].qout;
    g.rout = J->A[par].rout;
    g.accout = J->A[par].accout;
    g.s1lo = J->A[par].s1lo;
    g.s2lo = J->A[par].s2lo;
    g.qoutlo = J->A[par].qoutlo;
    g.accoutlo = J->A[par].accoutlo;
    g.fs = J->A[par].fs;
    return g;
}

template <typename R>
__device__ __forceinline__ int step_parity(cptr<Job<R>> J) {
    return as_const(J->ctl)->istep & 1;
}

// index of the job that owns work item w (jobs sorted by first)
template <typename R>
__device__ __forceinline__ int job_of(cptr<Job<R>> jobs, int njobs, int w) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Persistent launches: workgroup b takes items b, b + G, b + 2G, ... of the
// remapped list; with G a multiple of 8 one XCD works through a contiguous
// range of items (neighbouring tiles of one instance share halo planes in
// that XCD's L2).
template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool UNITV, bool K3>
__global__ __launch_bounds__((rhs_lanes<R, UNITV, TX, TY>()), (rhs_waves<R, UNITV>())) void k_rhs_batch(
    const Job<R> *__restrict__ jobs, int njobs, int nwork, int remap) {
    const cptr<Job<R>> cj = as_const(jobs);
    for (int w0 = blockIdx.x; w0 < nwork; w0 += gridDim.x) {
        const int w = remap_block(w0, nwork, remap);
        const cptr<Job<R>> J = cj + job_of(cj, njobs, w);
        const Params P = *(const Params *)&J->P;  // loads stay in the constant space
        if constexpr (rhs_split<R, UNITV>())
            rhs_tile_split<R, S, SV, TX, TY, WALLX, NONUNI, UNITV, K3>(P, stage_arrays(J, step_parity(J)),
                                                                      w - J->first);
        else
            rhs_tile<R, S, SV, TX, TY, WALLX, NONUNI, UNITV, K3>(P, stage_arrays(J, step_parity(J)), w - J->first);
        __syncthreads();  // LDS reuse by the next item
    }
}

template <typename R, int SV, int TX, int TY, bool WALLX, bool NONUNI>
__global__ __launch_bounds__(TX *TY) void k_dil_batch(const Job<R> *__restrict__ jobs, int njobs, int nwork,
                                                      int remap) {
    const cptr<Job<R>> cj = as_const(jobs);
    for (int w0 = blockIdx.x; w0 < nwork; w0 += gridDim.x) {
        const int w = remap_block(w0, nwork, remap);
        const cptr<Job<R>> J = cj + job_of(cj, njobs, w);
        const Params P = *(const Params *)&J->P;  // loads stay in the constant space
        dil_tile<R, SV, TX, TY, WALLX, NONUNI>(P, stage_arrays(J, step_parity(J)), J->dil, w - J->first);
        __syncthreads();
    }
}

// Start of a step, one lane per instance: on a CFL step whose viscous limit
// comes from the state, clear the viscous maximum first.
template <typename R>
__global__ void k_step_begin(const Job<R> *__restrict__ jobs, int njobs) {
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= njobs) return;
    const Job<R> &J = jobs[jb];
    StepCtl *c = J.ctl;
    if ((c->istep % c->check_cfl) == 0 && c->vfs) J.sc->visc_max = 0ull;
}

// CFL maxima of every instance due for a check; blockIdx.y is the instance
// (the body of k_cfl).
template <typename R>
__global__ void k_cfl_batch(const Job<R> *__restrict__ jobs) {
    const Job<R> &J = jobs[blockIdx.y];
    const StepCtl *c = J.ctl;
    if ((c->istep % c->check_cfl) != 0) return;
    const Params &P = J.P;
    const StagePtrs<R> &A = J.A[c->istep & 1];
    const int vfs = c->vfs;
    const long mxy = (long)P.mx * P.my, n = mxy * P.mz;
    double cmax = 0.0, vmax = 0.0;
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
        cfl_point(P, A, g, vfs, cmax, vmax);
    block_max2_to(cmax, vmax, &J.sc->conv_max, vfs ? &J.sc->visc_max : nullptr);
}

// Partial volume integrals for the forcing controller of every instance due
// for a check (the body of k_partials).
template <typename R>
__global__ void k_partials_batch(const Job<R> *__restrict__ jobs) {
    const Job<R> &J = jobs[blockIdx.y];
    const StepCtl *c = J.ctl;
    if (!c->forcing || (c->istep % c->check_cfl) != 0) return;
    partials_block(J.P, J.A[c->istep & 1], J.part);
}

// Per instance (one workgroup each): dt from the CFL maxima, the forcing
// controller, time += dt; then the viscous maximum is cleared for the last
// stage of this step, which accumulates it for the next check.
template <typename R>
__global__ void k_step_mid(const Job<R> *__restrict__ jobs) {
    const Job<R> &J = jobs[blockIdx.x];
    StepCtl *c = J.ctl;
    Scalars *sc = J.sc;
    const bool check = (c->istep % c->check_cfl) == 0;
    if (check && c->forcing) final_sums(J.part, sc, 1, 1, c->vol, nullptr, nullptr, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (check) {
            dt_finalize(c->cfl, sc);
            c->vfs = 0;
        }
        sc->time = sc->time + sc->dt;
        sc->visc_max = 0ull;
    }
}

template <typename R>
__global__ void k_step_end(const Job<R> *__restrict__ jobs, int njobs) {
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb < njobs) jobs[jb].ctl->istep += 1;
}

}  // namespace cns
