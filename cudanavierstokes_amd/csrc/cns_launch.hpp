// cns_launch.hpp -- host-side launchers of the templated kernels.  Each
// advective stencil half width S is compiled in its own translation unit
// (cns_kernels.hip built with -DCNS_S=1..4).
#pragma once
#include "cns_batch.hpp"

namespace cns {

// Tiles are TX x 16 columns.  The FP32 kernels take 32 x 16 tiles in
// 512-lane workgroups (384 halo points for 512 interior ones instead of 256
// for 256; measured 11-14% faster than 16 x 16 on order-6 channels, and the
// same at order 8); the FP64 role-split kernel 16 x 16 tiles with two lanes
// per column (512 lanes).
constexpr int TILE_Y = 16;
template <typename R, bool UNITV>
constexpr int tile_x() {
    return sizeof(R) == 4 ? 32 : 16;
}
inline int tile_x_of(bool fp32, bool unitv) { return fp32 ? 32 : 16; }

#define CNS_DECLARE_LAUNCH(N)                                                                                  \
    template <typename R>                                                                                      \
    hipError_t launch_rhs_s##N(int sv, bool wallx, bool nonuni, int nblocks, hipStream_t st, const Params &P, \
                               const StagePtrs<R> &A);                                                         \
    template <typename R>                                                                                      \
    hipError_t launch_dil_s##N(bool wallx, bool nonuni, int nblocks, hipStream_t st, const Params &P,         \
                               const StagePtrs<R> &A, R *dil);                                                 \
    template <typename R>                                                                                      \
    hipError_t launch_rhs_batch_s##N(int sv, bool wallx, bool nonuni, bool unitv, bool k3, int grid,           \
                                     hipStream_t st, const Job<R> *jobs, int njobs, int nwork, int remap);     \
    template <typename R>                                                                                      \
    hipError_t launch_dil_batch_s##N(bool wallx, bool nonuni, bool unitv, int grid, hipStream_t st,           \
                                     const Job<R> *jobs, int njobs, int nwork, int remap);
CNS_DECLARE_LAUNCH(1)
CNS_DECLARE_LAUNCH(2)
CNS_DECLARE_LAUNCH(3)
CNS_DECLARE_LAUNCH(4)
#undef CNS_DECLARE_LAUNCH

// boundary-layer forms (FP64, one solver; cns_kernels_bl.hip)
#define CNS_DECLARE_LAUNCH_BL(N)                                                                           \
    hipError_t launch_rhs_bl_s##N(int sv, bool nonuni, int nblocks, hipStream_t st, const Params &P,      \
                                  const StagePtrs<double> &A);                                           \
    hipError_t launch_dil_bl_s##N(bool nonuni, int nblocks, hipStream_t st, const Params &P,              \
                                  const StagePtrs<double> &A, double *dil);
CNS_DECLARE_LAUNCH_BL(1)
CNS_DECLARE_LAUNCH_BL(2)
CNS_DECLARE_LAUNCH_BL(3)
CNS_DECLARE_LAUNCH_BL(4)
#undef CNS_DECLARE_LAUNCH_BL

inline hipError_t launch_rhs_bl(int s, int sv, bool nonuni, int nblocks, hipStream_t st, const Params &P,
                                const StagePtrs<double> &A) {
    switch (s) {
    case 1: return launch_rhs_bl_s1(sv, nonuni, nblocks, st, P, A);
    case 2: return launch_rhs_bl_s2(sv, nonuni, nblocks, st, P, A);
    case 3: return launch_rhs_bl_s3(sv, nonuni, nblocks, st, P, A);
    case 4: return launch_rhs_bl_s4(sv, nonuni, nblocks, st, P, A);
    }
    return hipErrorInvalidValue;
}

inline hipError_t launch_dil_bl(int sv, bool nonuni, int nblocks, hipStream_t st, const Params &P,
                                const StagePtrs<double> &A, double *dil) {
    switch (sv) {
    case 1: return launch_dil_bl_s1(nonuni, nblocks, st, P, A, dil);
    case 2: return launch_dil_bl_s2(nonuni, nblocks, st, P, A, dil);
    case 3: return launch_dil_bl_s3(nonuni, nblocks, st, P, A, dil);
    case 4: return launch_dil_bl_s4(nonuni, nblocks, st, P, A, dil);
    }
    return hipErrorInvalidValue;
}

template <typename R>
inline hipError_t launch_rhs(int s, int sv, bool wallx, bool nonuni, int nblocks, hipStream_t st, const Params &P,
                             const StagePtrs<R> &A) {
    switch (s) {
    case 1: return launch_rhs_s1<R>(sv, wallx, nonuni, nblocks, st, P, A);
    case 2: return launch_rhs_s2<R>(sv, wallx, nonuni, nblocks, st, P, A);
    case 3: return launch_rhs_s3<R>(sv, wallx, nonuni, nblocks, st, P, A);
    case 4: return launch_rhs_s4<R>(sv, wallx, nonuni, nblocks, st, P, A);
    }
    return hipErrorInvalidValue;
}

// the dilatation kernel only depends on the viscous half width
template <typename R>
inline hipError_t launch_dil(int sv, bool wallx, bool nonuni, int nblocks, hipStream_t st, const Params &P,
                             const StagePtrs<R> &A, R *dil) {
    switch (sv) {
    case 1: return launch_dil_s1<R>(wallx, nonuni, nblocks, st, P, A, dil);
    case 2: return launch_dil_s2<R>(wallx, nonuni, nblocks, st, P, A, dil);
    case 3: return launch_dil_s3<R>(wallx, nonuni, nblocks, st, P, A, dil);
    case 4: return launch_dil_s4<R>(wallx, nonuni, nblocks, st, P, A, dil);
    }
    return hipErrorInvalidValue;
}

template <typename R>
inline hipError_t launch_rhs_batch(int s, int sv, bool wallx, bool nonuni, bool unitv, bool k3, int grid,
                                   hipStream_t st, const Job<R> *jobs, int njobs, int nwork, int remap) {
    switch (s) {
    case 1: return launch_rhs_batch_s1<R>(sv, wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
    case 2: return launch_rhs_batch_s2<R>(sv, wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
    case 3: return launch_rhs_batch_s3<R>(sv, wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
    case 4: return launch_rhs_batch_s4<R>(sv, wallx, nonuni, unitv, k3, grid, st, jobs, njobs, nwork, remap);
    }
    return hipErrorInvalidValue;
}

template <typename R>
inline hipError_t launch_dil_batch(int sv, bool wallx, bool nonuni, bool unitv, int grid, hipStream_t st,
                                   const Job<R> *jobs, int njobs, int nwork, int remap) {
    switch (sv) {
    case 1: return launch_dil_batch_s1<R>(wallx, nonuni, unitv, grid, st, jobs, njobs, nwork, remap);
    case 2: return launch_dil_batch_s2<R>(wallx, nonuni, unitv, grid, st, jobs, njobs, nwork, remap);
    case 3: return launch_dil_batch_s3<R>(wallx, nonuni, unitv, grid, st, jobs, njobs, nwork, remap);
    case 4: return launch_dil_batch_s4<R>(wallx, nonuni, unitv, grid, st, jobs, njobs, nwork, remap);
    }
    return hipErrorInvalidValue;
}

}  // namespace cns
