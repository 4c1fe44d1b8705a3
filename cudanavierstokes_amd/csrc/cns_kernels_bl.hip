// cns_kernels_bl.hip -- instantiates the boundary-layer forms of the RHS and
// dilatation kernels (FP64 storage, non-periodic x, one solver) for one
// advective half width S = CNS_S and every viscous half width SV <= S; the
// dilatation kernel for SV = CNS_S.
#include "cns_launch.hpp"
#include "cns_rhs.hpp"

#ifndef CNS_S
#error "compile with -DCNS_S=<1..4>"
#endif

namespace cns {

template <int S, int SV, bool UNITV, bool K3>
static hipError_t go_rhs_bl_v(bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<double> &A) {
    constexpr int TX = tile_x<double, UNITV>();
    const dim3 g(nb), b(rhs_lanes<double, UNITV, TX, TILE_Y>());
    if (nonuni)
        hipLaunchKernelGGL((k_rhs<double, S, SV, TX, TILE_Y, true, true, UNITV, K3, true>), g, b, 0, st, P, A);
    else
        hipLaunchKernelGGL((k_rhs<double, S, SV, TX, TILE_Y, true, false, UNITV, K3, true>), g, b, 0, st, P, A);
    return hipGetLastError();
}

template <int S, int SV>
static hipError_t go_rhs_bl(bool nonuni, int nb, hipStream_t st, const Params &P, const StagePtrs<double> &A) {
    const bool k3 = P.use_ra && P.use_rb;
    if (P.unit_visc)
        return k3 ? go_rhs_bl_v<S, SV, true, true>(nonuni, nb, st, P, A)
                  : go_rhs_bl_v<S, SV, true, false>(nonuni, nb, st, P, A);
    return k3 ? go_rhs_bl_v<S, SV, false, true>(nonuni, nb, st, P, A)
              : go_rhs_bl_v<S, SV, false, false>(nonuni, nb, st, P, A);
}

#define CNS_CAT2(a, b) a##b
#define CNS_CAT(a, b) CNS_CAT2(a, b)

hipError_t CNS_CAT(launch_rhs_bl_s, CNS_S)(int sv, bool nonuni, int nb, hipStream_t st, const Params &P,
                                          const StagePtrs<double> &A) {
    switch (sv) {
#if CNS_S >= 1
    case 1: return go_rhs_bl<CNS_S, 1>(nonuni, nb, st, P, A);
#endif
#if CNS_S >= 2
    case 2: return go_rhs_bl<CNS_S, 2>(nonuni, nb, st, P, A);
#endif
#if CNS_S >= 3
    case 3: return go_rhs_bl<CNS_S, 3>(nonuni, nb, st, P, A);
#endif
#if CNS_S >= 4
    case 4: return go_rhs_bl<CNS_S, 4>(nonuni, nb, st, P, A);
#endif
    }
    return hipErrorInvalidValue;
}

// here CNS_S is the viscous half width; the tiles are the RHS kernel's
// (16 x 16), whose last x tile holds the nodes the top ghosts extrapolate
hipError_t CNS_CAT(launch_dil_bl_s, CNS_S)(bool nonuni, int nb, hipStream_t st, const Params &P,
                                          const StagePtrs<double> &A, double *dil) {
    const dim3 g(nb), b(16 * TILE_Y);
    if (nonuni)
        hipLaunchKernelGGL((k_dil<double, CNS_S, 16, TILE_Y, true, true, true>), g, b, 0, st, P, A, dil);
    else
        hipLaunchKernelGGL((k_dil<double, CNS_S, 16, TILE_Y, true, false, true>), g, b, 0, st, P, A, dil);
    return hipGetLastError();
}

}  // namespace cns
