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
    int tile_x;  // x width of the launch's tiles (16 or 32; host dispatch of the dilatation kernel)
    // stage update, per conserved field:
    //   q_out   = s1 + dt (w_ra ra + w_rb rb + w_r rhs)   (when write_q)
    //   acc_out = s2 + dt (v_r rhs)                       (when write_acc)
    //   r_out   = rhs                                     (when write_r)
    int write_q, use_ra, use_rb, write_acc, write_r;
    int comp;       // FP32 storage: the state carries a compensation (low) part
    int write_qlo;  // FP32 storage: q_out gets its low part (the step result)
    double w_ra, w_rb, w_r, v_r;
    double d_dx, d_dy, d_dz, d_d2x, d_d2y, d_d2z;
    double Rgas, gam, cvInv, Re, Pr, Ec, viscexp, invPrEc, invRe, cpR;
    double twall_bot, twall_top;
    // FP32 storage keeps rho and rho E as offsets from a reference state
    // (rho0, E0) with T0 = p0/(rho0 R), h0 = (E0 + p0)/rho0, p0 = (gam-1) E0;
    // all four are zero for FP64 storage
    double rho0, E0, T0, h0;
    // FP32 storage, wall ghosts in offsets: wall temperature minus T0 at the
    // bottom and top walls, and cp T0 - h0 (zero up to rounding)
    double dtw_bot, dtw_top, ch0;
    // boundary layer (FP64): extrapolated x top and z ends, adiabatic x
    // bottom, optional wall-normal blowing at the bottom (omega1 = Re 121e-6,
    // lambda = Ly / 2 pi; time is the simulation time after the step's
    // advance, read on the device)
    int blayer, perturbed;
    double blow_om1, blow_lam;
    const double *time;
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
    const R *s1;   // state term of the update (stage input, step-start state or accumulator)
    const R *ra;   // previous rhs (low-storage RK3 r_prev; Kutta k1)
    const R *rb;   // Kutta k2
    const R *s2;   // state term of the accumulator update (RK4)
    R *qout;       // updated state
    R *rout;       // right-hand side output
    R *accout;     // accumulator output (RK4)
    // FP32 storage: low parts of the compensated state (x = hi + lo), one per
    // state array above (unused with FP64 storage)
    const R *s1lo, *s2lo;
    R *qoutlo, *accoutlo;
    long fs;  // field stride (points per field)
};

// FP32 storage computes with offsets from the reference state
template <typename R>
struct Off {
    static constexpr bool on = sizeof(R) == 4;
};

// Central-difference weights of order 2s as compile-time constants (exact
// rationals, one rounding each, so they equal the literal tables of
// src/globals.h:69-95 bitwise):
//   D1 f = sum_{k=1..s} a_k (f[i+k] - f[i-k]),  a_k = (-1)^{k+1} (s!)^2 / (k (s-k)! (s+k)!)
//   D2 f = b_0 f[i] + sum_{k=1..s} b_k (f[i+k] + f[i-k]),
//          b_k = 2 (-1)^{k+1} (s!)^2 / (k^2 (s-k)! (s+k)!),  b_0 = -2 sum b_k
// The reference's coeffF[it] is -a_{s-it} and coeffS[it] is b_{s-it}.
constexpr long long c_fact(int n) {
    long long f = 1;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}
constexpr long long c_gcd(long long a, long long b) {
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    while (b) {
        long long t = a % b;
        a = b;
        b = t;
    }
    return a ? a : 1;
}
constexpr double fd1(int s, int k) {
    long long num = c_fact(s) * c_fact(s), den = (long long)k * c_fact(s - k) * c_fact(s + k);
    const long long g = c_gcd(num, den);
    num /= g;
    den /= g;
    return ((k % 2) ? (double)num : -(double)num) / (double)den;
}
constexpr double fd2(int s, int k) {
    if (k > 0) {
        long long num = 2 * c_fact(s) * c_fact(s), den = (long long)k * k * c_fact(s - k) * c_fact(s + k);
        const long long g = c_gcd(num, den);
        num /= g;
        den /= g;
        return ((k % 2) ? (double)num : -(double)num) / (double)den;
    }
    long long cn = 0, cd = 1;  // running sum of b_k as a reduced fraction
    for (int j = 1; j <= s; j++) {
        long long num = 2 * c_fact(s) * c_fact(s), den = (long long)j * j * c_fact(s - j) * c_fact(s + j);
        long long g = c_gcd(num, den);
        num /= g;
        den /= g;
        if (!(j % 2)) num = -num;
        long long nn = cn * den + num * cd, dn = cd * den;
        g = c_gcd(nn, dn);
        cn = nn / g;
        cd = dn / g;
    }
    long long nc = -2 * cn, dc = cd;
    const long long g = c_gcd(nc, dc);
    return (double)(nc / g) / (double)(dc / g);
}

// The weights of one order, evaluated in constant initialisers so that no
// arithmetic of fd1/fd2 survives into device code.
template <int S>
struct FD {
    static constexpr double A1 = fd1(S, 1), A2 = S >= 2 ? fd1(S, 2) : 0.0, A3 = S >= 3 ? fd1(S, 3) : 0.0,
                            A4 = S >= 4 ? fd1(S, 4) : 0.0;
    static constexpr double B0 = fd2(S, 0), B1 = fd2(S, 1), B2 = S >= 2 ? fd2(S, 2) : 0.0,
                            B3 = S >= 3 ? fd2(S, 3) : 0.0, B4 = S >= 4 ? fd2(S, 4) : 0.0;
    __host__ __device__ static constexpr double a(int k) { return k == 1 ? A1 : k == 2 ? A2 : k == 3 ? A3 : A4; }
    __host__ __device__ static constexpr double b(int k) {
        return k == 0 ? B0 : k == 1 ? B1 : k == 2 ? B2 : k == 3 ? B3 : B4;
    }
};

__device__ __forceinline__ int wrapi(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}
// wrapi for a in [-n, 2n), without the division: z-plane indices within a
// stencil half width of the domain (n > 2S holds for every grid)
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

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

// T^n of the FP64 viscosity law mu = T^n / Re (T > 0): T sqrt(T) for the
// boundary layer's n = 3/2, else exp2(n log2 T); both within a few ulp of
// pow(T, n) at 7% and 60% of its instructions (pow handles every sign and
// special case of its arguments, which a temperature never needs).
__device__ __forceinline__ double visc_pow(double t, double n) { return n == 1.5 ? t * sqrt(t) : exp2(n * log2(t)); }

// Equation of state and viscosity law (src/cuda_main.cu:228-239).
template <typename R, bool UNITV>
__device__ __forceinline__ void eos(const Params &P, R r, R ru, R rv, R rw, R re, R &u, R &v, R &w, R &t,
                                    R &h, R &mu) {
    R inv = R(1) / r;
    u = ru * inv;
    v = rv * inv;
    w = rw * inv;
    const R k2 = R(0.5) * (u * u + v * v + w * w);
    R en = re * inv - k2;
    t = R(P.cvInv) * en;
    // total enthalpy (rhoE + p)/rho = cp T + |u|^2/2 with cp = gam R/(gam-1)
    h = t * R(P.cpR) + k2;
    // mu = T^viscexp / Re; for viscexp = 1 the kernels rebuild it as T/Re
    mu = UNITV ? R(0) : R(visc_pow((double)t, P.viscexp)) * R(P.invRe);
}

// Isothermal no-slip wall ghost from its mirror point (src/boundary.h:79-148):
// velocity antisymmetric, pressure and dilatation symmetric, T = 2Tw - T,
// mu and lambda from the ghost T, rho = p/(R T), h = cp T + |u|^2/2.
template <typename R, bool UNITV>
__device__ __forceinline__ void wall_ghost(const Params &P, R tw, R &r, R &u, R &v, R &w, R &t, R &h, R &mu) {
    R p = r * R(P.Rgas) * t;
    u = -u;
    v = -v;
    w = -w;
    t = R(2.0) * tw - t;
    mu = UNITV ? R(0) : R(visc_pow((double)t, P.viscexp)) * R(P.invRe);
    h = t * R(P.cpR) + R(0.5) * (u * u + v * v + w * w);
    r = p / (R(P.Rgas) * t);
}

// Boundary layer, bottom x boundary (src/boundary_condition_x.h:60-74,
// src/boundary.h:121-180): an adiabatic no-slip reflection about the face.
// The ghost is the mirrored point with odd velocities and even p and T (so
// T and mu are kept), u replaced by the blowing value ub when blow is set,
// and rho and h formed from the ghost p, T and velocities.
template <typename R>
__device__ __forceinline__ void bl_bottom_ghost(const Params &P, bool blow, R ub, R &r, R &u, R &v, R &w, R t,
                                                R &h) {
    const R p = r * R(P.Rgas) * t;
    u = blow ? ub : -u;
    v = -v;
    w = -w;
    h = t * R(P.cpR) + R(0.5) * (u * u + v * v + w * w);
    r = p / (R(P.Rgas) * t);
}

// Wall-normal blowing of src/perturbation.h:25-53 at global z index k: on
// k in [90, 130] it is kappa g(xi) [2e-4 s1 + 3e-5 s2 cos(y/lambda)] with
// g = 15.1875 xi^5 - 35.4375 xi^4 + 20.25 xi^3, kappa = +1 below k = 110 and
// -1 from there on, xi the integer distance from the nearer end of the band
// over 20; s1 = sin(omega1 t), s2 = sin(10 t) and cosj = cos(y/lambda) come
// from the caller.  Returns whether k is in the band.
__device__ __forceinline__ bool bl_blow(int k, double s1, double s2, double cosj, double &val) {
    const int kC = 110, LP = 40, kSt = kC - LP / 2, kEn = kC + LP / 2;
    if (k < kSt || k > kEn) return false;
    const int kappa = k < kC ? 1 : -1;
    const int alpha = k < kC ? k - kSt : kEn - k;
    const int beta = k < kC ? kC - kSt : kEn - kC;
    const double xi = alpha * 1.0 / beta;
    const double g = (15.1875 * xi * xi * xi * xi * xi) - (35.4375 * xi * xi * xi * xi) + (20.25 * xi * xi * xi);
    val = 2e-4 * kappa * g * s1 + 3e-5 * kappa * g * s2 * cosj;
    return true;
}

// 1/x: correctly rounded division for FP64 storage (parity with the FP64
// reference); the hardware reciprocal (1 ulp) for FP32 storage, whose
// tolerance is 1e-5.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Viscosity law mu = T^n / Re for FP32 storage: 2^(n log2 T) with the
// hardware base-2 logarithm and exponential (about 1 ulp each, so mu is
// accurate to a few parts in 10^7, inside the 1e-5 budget of FP32 storage);
// the FP64 path keeps pow.
__device__ __forceinline__ float visc_law(float t, float n, float inv_re) {
    return __builtin_amdgcn_exp2f(n * __builtin_amdgcn_logf(t)) * inv_re;
}

// Equation of state on offset storage (FP32): from rho' = rho - rho0 and
// E' = rhoE - E0,
//   p' = (gam-1)(E' - rho |u|^2/2),  T' = (p' - rho' R T0)/(rho R),
//   h' = (E' + p' - rho' h0)/rho,
// so that T = T0 + T' and h = h0 + h' carry the fluctuation at full relative
// precision (p0, T0, h0 cancel exactly in the algebra).
template <typename R, bool UNITV>
__device__ __forceinline__ void eos_off(const Params &P, R rp, R ru, R rv, R rw, R ep, R &u, R &v, R &w, R &tp,
                                        R &hp, R &mu) {
    const R r = R(P.rho0) + rp;
    const R inv = recip(r);
    u = ru * inv;
    v = rv * inv;
    w = rw * inv;
    const R k2 = R(0.5) * (u * u + v * v + w * w);
    const R pp = R(P.gam - 1.0) * (ep - r * k2);
    tp = (pp - rp * R(P.Rgas * P.T0)) * inv * R(1.0 / P.Rgas);
    hp = (ep + pp - rp * R(P.h0)) * inv;
    mu = UNITV ? R(0) : visc_law(R(P.T0) + tp, R(P.viscexp), R(P.invRe));
}

// Wall ghost on offset storage (FP32), from the mirror point's offsets:
// with T_g = 2 Tw - T and p_g = p, rho_g = rho T / T_g, so
//   T'_g   = 2 (Tw - T0) - T',
//   rho'_g = (2 rho0 (T0 - Tw) + 2 rho0 T' + rho' T) / T_g,
//   h'_g   = cp T'_g + |u|^2/2 + (cp T0 - h0),
// every term of the size of the fluctuation, so FP32 keeps its relative
// precision (dtw = Tw - T0 in FP64 on the host).
template <typename R, bool UNITV>
__device__ __forceinline__ void wall_ghost_off(const Params &P, R dtw, R &rp, R &u, R &v, R &w, R &tp, R &hp,
                                               R &mu) {
    const R T0 = R(P.T0), r0 = R(P.rho0);
    const R t = T0 + tp;
    const R tpg = R(2) * dtw - tp;
    const R tg = T0 + tpg;
    rp = (R(2) * r0 * (tp - dtw) + rp * t) / tg;
    u = -u;
    v = -v;
    w = -w;
    tp = tpg;
    hp = R(P.cpR) * tpg + R(0.5) * (u * u + v * v + w * w) + R(P.ch0);
    mu = UNITV ? R(0) : visc_law(tg, R(P.viscexp), R(P.invRe));
}

// Quantity indices of the staged primitives (LDS plane and z register
// window).  For mu = T/Re the last one is not staged: it is rebuilt from T.
enum { QR = 0, QU, QV, QW, QH, QT, QD, QM, NQ };
// pressure over nothing else: the boundary layer's z window carries p = rho R T
// as its own quantity, since the reference extrapolates it on its own
constexpr int QP = NQ;

}  // namespace cns
