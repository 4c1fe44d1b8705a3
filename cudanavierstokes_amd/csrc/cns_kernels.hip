// cns_kernels.hip -- instantiates the RHS and dilatation kernels (single
// instance and batched) for one advective half width S = CNS_S, every
// viscous half width SV <= S and both x-boundary / x-grid variants; for one
// storage type CNS_R (double or float), or both when CNS_R is not defined.
#include "cns_batch.hpp"
#include "cns_launch.hpp"
#include "cns_rhs.hpp"

#ifndef CNS_S
#error "compile with -DCNS_S=<1..4>"
#endif

namespace cns {

template <typename R, int S, int SV, bool UNITV, bool K3>
static hipError_t go_rhs_v(bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<R> &A) {
    constexpr int TX = tile_x<R, UNITV>();
    const dim3 g(nb), b(rhs_lanes<R, UNITV, TX, TILE_Y>());
    if (wallx) {
        if (nonuni)
            hipLaunchKernelGGL((k_rhs<R, S, SV, TX, TILE_Y, true, true, UNITV, K3>), g, b, 0, st, P, A);
        else
            hipLaunchKernelGGL((k_rhs<R, S, SV, TX, TILE_Y, true, false, UNITV, K3>), g, b, 0, st, P, A);
    } else {
        if (nonuni)
            hipLaunchKernelGGL((k_rhs<R, S, SV, TX, TILE_Y, false, true, UNITV, K3>), g, b, 0, st, P, A);
        else
            hipLaunchKernelGGL((k_rhs<R, S, SV, TX, TILE_Y, false, false, UNITV, K3>), g, b, 0, st, P, A);
    }
    return hipGetLastError();
}

template <typename R, int S, int SV>
static hipError_t go_rhs(bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<R> &A) {
    // three update operands only in the last Kutta stage (q0, k1, k2)
    const bool k3 = P.use_ra && P.use_rb;
    if (P.unit_visc)
        return k3 ? go_rhs_v<R, S, SV, true, true>(wallx, nonuni, nb, st, P, A)
                  : go_rhs_v<R, S, SV, true, false>(wallx, nonuni, nb, st, P, A);
    return k3 ? go_rhs_v<R, S, SV, false, true>(wallx, nonuni, nb, st, P, A)
              : go_rhs_v<R, S, SV, false, false>(wallx, nonuni, nb, st, P, A);
}

template <typename R, int SV, int TX>
static hipError_t go_dil_t(bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<R> &A,
                           R *dil) {
    const dim3 g(nb), b(TX * TILE_Y);
    if (wallx) {
        if (nonuni)
            hipLaunchKernelGGL((k_dil<R, SV, TX, TILE_Y, true, true>), g, b, 0, st, P, A, dil);
        else
            hipLaunchKernelGGL((k_dil<R, SV, TX, TILE_Y, true, false>), g, b, 0, st, P, A, dil);
    } else {
        if (nonuni)
            hipLaunchKernelGGL((k_dil<R, SV, TX, TILE_Y, false, true>), g, b, 0, st, P, A, dil);
        else
            hipLaunchKernelGGL((k_dil<R, SV, TX, TILE_Y, false, false>), g, b, 0, st, P, A, dil);
    }
    return hipGetLastError();
}

// tile width of the launch as the host set it (P.tile_x)
template <typename R, int SV>
static hipError_t go_dil(bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<R> &A,
                         R *dil) {
    if (P.tile_x == 32) return go_dil_t<R, SV, 32>(wallx, nonuni, nb, st, P, A, dil);
    return go_dil_t<R, SV, 16>(wallx, nonuni, nb, st, P, A, dil);
}

template <typename R, int S, int SV, bool UNITV, bool K3>
static hipError_t go_rhs_batch_v(bool wallx, bool nonuni, int grid, hipStream_t st, const Job<R> *jobs, int njobs,
                                 int nwork, int remap) {
    constexpr int TX = tile_x<R, UNITV>();
    const dim3 g(grid), b(rhs_lanes<R, UNITV, TX, TILE_Y>());
    if (wallx) {
        if (nonuni)
            hipLaunchKernelGGL((k_rhs_batch<R, S, SV, TX, TILE_Y, true, true, UNITV, K3>), g, b, 0, st, jobs, njobs,
                               nwork, remap);
        else
            hipLaunchKernelGGL((k_rhs_batch<R, S, SV, TX, TILE_Y, true, false, UNITV, K3>), g, b, 0, st, jobs, njobs,
                               nwork, remap);
    } else {
        if (nonuni)
            hipLaunchKernelGGL((k_rhs_batch<R, S, SV, TX, TILE_Y, false, true, UNITV, K3>), g, b, 0, st, jobs, njobs,
                               nwork, remap);
        else
            hipLaunchKernelGGL((k_rhs_batch<R, S, SV, TX, TILE_Y, false, false, UNITV, K3>), g, b, 0, st, jobs,
                               njobs, nwork, remap);
    }
    return hipGetLastError();
}

template <typename R, int S, int SV>
static hipError_t go_rhs_batch(bool wallx, bool nonuni, bool unitv, bool k3, int grid, hipStream_t st,
                               const Job<R> *jobs, int njobs, int nwork, int remap) {
    if (unitv)
        return k3 ? go_rhs_batch_v<R, S, SV, true, true>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap)
                  : go_rhs_batch_v<R, S, SV, true, false>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap);
    return k3 ? go_rhs_batch_v<R, S, SV, false, true>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap)
              : go_rhs_batch_v<R, S, SV, false, false>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap);
}

template <typename R, int SV, int TX>
static hipError_t go_dil_batch_t(bool wallx, bool nonuni, int grid, hipStream_t st, const Job<R> *jobs, int njobs,
                                 int nwork, int remap) {
    const dim3 g(grid), b(TX * TILE_Y);
    if (wallx) {
        if (nonuni)
            hipLaunchKernelGGL((k_dil_batch<R, SV, TX, TILE_Y, true, true>), g, b, 0, st, jobs, njobs, nwork,
                               remap);
        else
            hipLaunchKernelGGL((k_dil_batch<R, SV, TX, TILE_Y, true, false>), g, b, 0, st, jobs, njobs, nwork,
                               remap);
    } else {
        if (nonuni)
            hipLaunchKernelGGL((k_dil_batch<R, SV, TX, TILE_Y, false, true>), g, b, 0, st, jobs, njobs, nwork,
                               remap);
        else
            hipLaunchKernelGGL((k_dil_batch<R, SV, TX, TILE_Y, false, false>), g, b, 0, st, jobs, njobs, nwork,
                               remap);
    }
    return hipGetLastError();
}

template <typename R, int SV>
static hipError_t go_dil_batch(bool wallx, bool nonuni, bool unitv, int grid, hipStream_t st, const Job<R> *jobs,
                               int njobs, int nwork, int remap) {
    if (tile_x_of(sizeof(R) == 4, unitv) == 32)
        return go_dil_batch_t<R, SV, tile_x<R, true>()>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap);
    return go_dil_batch_t<R, SV, 16>(wallx, nonuni, grid, st, jobs, njobs, nwork, remap);
}

#define CNS_CAT2(a, b) a##b
#define CNS_CAT(a, b) CNS_CAT2(a, b)
#define LAUNCH_RHS CNS_CAT(launch_rhs_s, CNS_S)
#define LAUNCH_DIL CNS_CAT(launch_dil_s, CNS_S)
#define LAUNCH_RHS_BATCH CNS_CAT(launch_rhs_batch_s, CNS_S)
#define LAUNCH_DIL_BATCH CNS_CAT(launch_dil_batch_s, CNS_S)

template <typename R>
hipError_t LAUNCH_RHS(int sv, bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P,
                      const StagePtrs<R> &A) {
    switch (sv) {
#if CNS_S >= 1
    case 1: return go_rhs<R, CNS_S, 1>(wallx, nonuni, nb, st, P, A);
#endif
#if CNS_S >= 2
    case 2: return go_rhs<R, CNS_S, 2>(wallx, nonuni, nb, st, P, A);
#endif
#if CNS_S >= 3
    case 3: return go_rhs<R, CNS_S, 3>(wallx, nonuni, nb, st, P, A);
#endif
#if CNS_S >= 4
    case 4: return go_rhs<R, CNS_S, 4>(wallx, nonuni, nb, st, P, A);
#endif
    }
    return hipErrorInvalidValue;
}

// here CNS_S is the viscous half width of the dilatation kernel
template <typename R>
hipError_t LAUNCH_DIL(bool wallx, bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<R> &A,
                      R *dil) {
    return go_dil<R, CNS_S>(wallx, nonuni, nb, st, P, A, dil);
}

template <typename R>
hipError_t LAUNCH_RHS_BATCH(int sv, bool wallx, bool nonuni, bool unitv, bool k3, int grid, hipStream_t st,
                            const Job<R> *jobs, int njobs, int nwork, int remap) {
    switch (sv) {
#if CNS_S >= 1
    case 1: return go_rhs_batch<R, CNS_S, 1>(wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
#endif
#if CNS_S >= 2
    case 2: return go_rhs_batch<R, CNS_S, 2>(wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
#endif
#if CNS_S >= 3
    case 3: return go_rhs_batch<R, CNS_S, 3>(wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
#endif
#if CNS_S >= 4
    case 4: return go_rhs_batch<R, CNS_S, 4>(wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
#endif
    }
    return hipErrorInvalidValue;
}

// here CNS_S is the viscous half width of the dilatation kernel
template <typename R>
hipError_t LAUNCH_DIL_BATCH(bool wallx, bool nonuni, bool unitv, int grid, hipStream_t st, const Job<R> *jobs,
                            int njobs, int nwork, int remap) {
    return go_dil_batch<R, CNS_S>(wallx, nonuni, unitv, grid, st, jobs, njobs, nwork, remap);
}

#define CNS_INSTANTIATE(R)                                                                                        \
    template hipError_t LAUNCH_RHS<R>(int, bool, bool, int, hipStream_t, const Params &, const StagePtrs<R> &);  \
    template hipError_t LAUNCH_DIL<R>(bool, bool, int, hipStream_t, const Params &, const StagePtrs<R> &, R *);  \
    template hipError_t LAUNCH_RHS_BATCH<R>(int, bool, bool, bool, bool, int, hipStream_t, const Job<R> *, int, int, \
                                            int);                                                                 \
    template hipError_t LAUNCH_DIL_BATCH<R>(bool, bool, bool, int, hipStream_t, const Job<R> *, int, int, int);

#if defined(CNS_R)
CNS_INSTANTIATE(CNS_R)
#else
CNS_INSTANTIATE(double)
CNS_INSTANTIATE(float)
#endif

}  // namespace cns
