// cns_device.hpp -- gfx950 kernels of the explicit compressible Navier-Stokes
// time advance (right-hand side, ghost fills, sponge, RK stage updates,
// CFL/bulk reductions).
//
// Data layout: five conserved fields (rho, rho*u, rho*v, rho*w, rho*E), one
// array each, idx = i + mx*(j + my*k) with x unit-stride (the layout of
// src/globals.h:60-61).  The RHS kernel is a 2.5D sweep: a workgroup owns a
// TX x TY tile of x-y columns and marches along z.  The current output plane
// of the tile plus an x/y halo of width S lives in LDS (x and y stencils);
// each lane keeps a (2S+1)-plane window of its own column in registers
// (z stencils).  Ghost values (periodic wrap, isothermal no-slip wall) are
// produced while staging, so no separate boundary kernel exists.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cns {

// Per-launch parameters (passed by value as a kernel argument).
struct Params {
    int mx, my, mz;
    int zc;             // planes per z chunk
    int ntx, nty, nzc;  // tiles along x, y and chunks along z
    int forcing, sponge, fix_y_work, unit_visc;
    int want_visc_max;  // accumulate max(mu/dx^2, ...) for the next CFL check
    int xcd_remap;
    // stage update, one linear combination per conserved field:
    //   q_out = w_qc q_in + w_q0 q0 + w_acc acc + dt (w_ra ra + w_rb rb + w_r rhs)
    //   acc   = v_qc q_in + v_acc acc + dt (v_r rhs)        (when write_acc)
    //   r_out = rhs                                          (when write_r)
    int use_q0, use_acc, use_ra, use_rb, write_q, write_r, write_acc;
    double w_qc, w_q0, w_acc, w_ra, w_rb, w_r, v_qc, v_acc, v_r;
    double d_dx, d_dy, d_dz, d_d2x, d_d2y, d_d2z;
    double Rgas, gam, cvInv, Re, Pr, Ec, viscexp, invPrEc;
    double twall_bot, twall_top;
    double cF[4], cS[5], cVF[4], cVS[5];
    const double *xp, *vsx, *dxv, *spx, *spz;
    const double *ref;  // 5 sponge reference fields of mx*mz, contiguous
    const double *dt, *dpdz;
    unsigned long long *visc_max_out;
};

// Field sets: the five conserved fields (rho, rho u, rho v, rho w, rho E) of
// one state are one allocation, field q at offset q*fs.
template <typename R>
struct StagePtrs {
    const R *qin;  // stage input (read with stencils)
    const R *dil;  // dilatation of qin (one field)
    R *qout;       // updated state
    R *rout;       // right-hand side output
    const R *ra;   // previous rhs (low-storage RK3 r_prev; Kutta k1)
    const R *rb;   // Kutta k2
    const R *q0;   // step-start state (Kutta, RK4)
    R *acc;        // RK4 accumulator
    long fs;       // field stride (points per field)
};

__device__ __forceinline__ int wrapi(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

// Consecutive workgroup ids are spread round-robin over the 8 XCDs; remap so
// that each XCD receives a contiguous range of tiles (neighbouring tiles share
// halos through that XCD's L2).
__device__ __forceinline__ int remap_block(int b, int nb, int on) {
    if (!on || nb < 8) return b;
    int per = nb / 8, rem = nb % 8;
    int xcd = b % 8, slot = b / 8;
    // XCDs 0..rem-1 hold per+1 blocks, the rest hold per
    int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return start + slot;
}

// Equation of state and viscosity law (src/cuda_main.cu:228-239).
template <typename R>
__device__ __forceinline__ void eos(const Params &P, R r, R ru, R rv, R rw, R re, R &u, R &v, R &w, R &t,
                                    R &h, R &mu) {
    R inv = R(1) / r;
    u = ru * inv;
    v = rv * inv;
    w = rw * inv;
    R en = re * inv - R(0.5) * (u * u + v * v + w * w);
    t = R(P.cvInv) * en;
    R p = r * R(P.Rgas) * t;
    h = (re + p) * inv;
    R suth = P.unit_visc ? t : R(pow((double)t, P.viscexp));
    mu = suth / R(P.Re);
}

// Isothermal no-slip wall ghost from its mirror point (src/boundary.h:79-148):
// velocity antisymmetric, pressure and dilatation symmetric, T = 2Tw - T,
// mu and lambda from the ghost T, rho = p/(R T), h = cp T + |u|^2/2.
template <typename R>
__device__ __forceinline__ void wall_ghost(const Params &P, R tw, R &r, R &u, R &v, R &w, R &t, R &h, R &mu) {
    R p = r * R(P.Rgas) * t;
    u = -u;
    v = -v;
    w = -w;
    t = R(2.0) * tw - t;
    R suth = P.unit_visc ? t : R(pow((double)t, P.viscexp));
    mu = suth / R(P.Re);
    h = t * R(P.Rgas) * R(P.gam) / (R(P.gam) - R(1.0)) + R(0.5) * (u * u + v * v + w * w);
    r = p / (R(P.Rgas) * t);
}

// Quantity indices of the staged arrays.
enum { QR = 0, QU, QV, QW, QH, QT, QM, QD, NQ };

}  // namespace cns
