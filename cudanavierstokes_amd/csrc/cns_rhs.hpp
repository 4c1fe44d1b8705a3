// cns_rhs.hpp -- fused right-hand side + RK stage update and dilatation
// kernels (templates; instantiated per stencil order in cns_kernels_*.hip).
//
// One RHS evaluation follows the reference term by term:
//   src/cuda_main.cu:15-42   calcRHS: calcState -> derVel -> calcDil ->
//                            RHSX (=) -> RHSY (+=) -> RHSZ (+=) -> addSponge
//   src/cuda_rhs.cu:9-128    x part     src/cuda_rhs.cu:130-262  y part
//   src/cuda_rhs.cu:264-396  z part with the dp/dz forcing
//   src/cuda_derivs.h:30-312 stencils and split-form fluxes
// but everything between reading the state and writing the update happens in
// registers and LDS: no T/p/h/mu/lambda or velocity-gradient fields exist.
//
// Split-form flux.  The reference evaluates, per point, the interface fluxes
// F(i+1/2) and F(i-1/2) as double sums over (l, m) pairs
// (src/cuda_derivs.h:30-155).  Their difference telescopes within each l:
//   F(i+1/2) - F(i-1/2) = sum_l c_l [P_l(i) - P_l(i-l)],
//   P_l(j) = (f_j + f_{j+l})(g_j + g_{j+l})(h_j + h_{j+l}),
// which is the same discrete operator with 2S pair products per point
// instead of S(S+1).
#pragma once
#include "cns_device.hpp"

#include <type_traits>

namespace cns {

// Source of one staged point: the in-plane byte offset of the point whose
// value is used (y wraps; x wraps, or mirrors at a wall) and the wall side
// (-1 bottom, +1 top, 0 none) whose ghost rule then applies
// (src/boundary_condition_x.h, src/boundary.h:79-148).
template <bool WALLX>
__device__ __forceinline__ unsigned src_offset(const Params &P, int gi, int gj, unsigned esz, int &side) {
    gj = wrapi(gj, P.my);
    side = 0;
    if (WALLX) {
        if (gi < 0) {
            side = -1;
            gi = -1 - gi;
        } else if (gi >= P.mx) {
            side = 1;
            gi = 2 * P.mx - 1 - gi;
        }
        gi = gi < 0 ? 0 : (gi >= P.mx ? P.mx - 1 : gi);
    } else {
        gi = wrapi(gi, P.mx);
    }
    return (unsigned)(gi + P.mx * gj) * esz;
}

// Field accesses are raw buffer loads and stores: a uniform plane base in a
// buffer resource (scalar registers, rebuilt by scalar instructions per
// plane) plus a 32-bit per-lane byte offset, so no 64-bit address arithmetic
// runs on the vector ALU.  The resource covers 2^32 bytes from the base with
// no stride (third word 0x00020000 for gfx9 raw buffers).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}
using u32x2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc_of(nullptr), 0, 0, 0));
__device__ __forceinline__ double ld_at(const double *base, unsigned boff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc_of(base), (int)boff, 0, 0));
}
__device__ __forceinline__ float ld_at(const float *base, unsigned boff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc_of(base), (int)boff, 0, 0));
}
__device__ __forceinline__ void st_at(double *base, unsigned boff, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rsrc_of(base), (int)boff, 0, 0);
}
__device__ __forceinline__ void st_at(float *base, unsigned boff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rsrc_of(base), (int)boff, 0, 0);
}

// Conserved values and dilatation of one point, as loaded.
template <typename R>
struct Raw {
    R c[5], d;
};

template <typename R>
__device__ __forceinline__ void load_raw(const StagePtrs<R> &A, long poff, unsigned boff, Raw<R> &x) {
    const R *b = A.qin + poff;
#pragma unroll
    for (int q = 0; q < 5; q++) x.c[q] = ld_at(b + q * A.fs, boff);
    x.d = ld_at(A.dil + poff, boff);
}

// Primitives of one loaded point with the x ghost rule of its side applied.
// On offset storage (FP32) rho, T and h are staged as rho', T', h'.
template <typename R, bool WALLX, bool UNITV>
__device__ __forceinline__ void to_prim(const Params &P, const Raw<R> &x, int side, R *q) {
    R r = x.c[0], u, v, w, t, h, mu;
    if constexpr (Off<R>::on) {
        eos_off<R, UNITV>(P, r, x.c[1], x.c[2], x.c[3], x.c[4], u, v, w, t, h, mu);
        if (WALLX && side != 0) wall_ghost_off<R, UNITV>(P, R(side < 0 ? P.dtw_bot : P.dtw_top), r, u, v, w, t, h, mu);
    } else {
        eos<R, UNITV>(P, r, x.c[1], x.c[2], x.c[3], x.c[4], u, v, w, t, h, mu);
        if (WALLX && side != 0)
            wall_ghost<R, UNITV>(P, side < 0 ? R(P.twall_bot) : R(P.twall_top), r, u, v, w, t, h, mu);
    }
    q[QR] = r;
    q[QU] = u;
    q[QV] = v;
    q[QW] = w;
    q[QH] = h;
    q[QT] = t;
    q[QD] = x.d;
    q[QM] = mu;
}

// Viscous-order first and second derivatives of u, v, w along DIR
// (src/cuda_derivs.h:192-224 for x, 248-268 for y, 292-312 for z).  The
// weights are compile-time constants; the spacing multiplies each sum once.
template <typename R, int SV, int DIR, bool NONUNI, class Acc>
__device__ __forceinline__ void vel_stencils(const Params &P, Acc f, const double *vs, double xpi, R d1[3],
                                             R d2[3]) {
    const R d = R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    const R dd = R(DIR == 0 ? P.d_d2x : (DIR == 1 ? P.d_d2y : P.d_d2z));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int q = QU + c;
        R a1 = R(0), a2 = R(0);
#pragma unroll
        for (int k = SV; k >= 1; k--) {
            const R fp = f(q, k), fm = f(q, -k);
            a1 += R(FD<SV>::a(k)) * (fp - fm);
            a2 += R(FD<SV>::b(k)) * (fp + fm);
        }
        if (Off<R>::on && DIR == 0 && NONUNI) {
            // stretched-grid metric terms in FP64
            d1[c] = R((double)a1 * (P.d_dx * xpi));
        } else {
            a1 = a1 * d;
            if (DIR == 0 && NONUNI) a1 = a1 * R(xpi);
            d1[c] = a1;
        }
        if (DIR == 0 && NONUNI) {
            if constexpr (Off<R>::on) {
                double b = 0.0;
#pragma unroll
                for (int it = 0; it < 2 * SV + 1; it++) b += vs[it] * (double)f(q, it - SV);
                d2[c] = R(b);
            } else {
                R b = R(0);
#pragma unroll
                for (int it = 0; it < 2 * SV + 1; it++) b += R(vs[it]) * f(q, it - SV);
                d2[c] = b;
            }
        } else {
            d2[c] = (a2 + R(FD<SV>::b(0)) * f(q, 0)) * dd;
        }
    }
}

// The part of one direction's RHS that does not involve the velocity
// gradients (fluxes, pressure gradient, heat-flux and dilatation
// derivatives), kept for the stress part below.
template <typename R>
struct DirA {
    R uc, vc, wc, muc, dilc, lamc, d1m, d1t, d1d, d2t, d1p, d1l, cr, cu, cv, cw, ce;
};

// Split-form convective fluxes of one direction (the telescoped pair form):
// -d(fgh)/dx ~ -(d/4) sum_l a_l [(f_i+f_{i+l})(g_i+g_{i+l})(h_i+h_{i+l}) - (f_{i-l}+f_i)(...)(...)],
// out = mass, three momentum and energy fluxes; xs is the direction's
// inverse spacing (with the x metric).
template <typename R, int S, int DIR, class Acc>
__device__ __forceinline__ void dir_flux(const Params &P, Acc f, R xs, R out[5]) {
    const int QA = (DIR == 0) ? QU : (DIR == 1 ? QV : QW);
    R fr = R(0), fu = R(0), fv = R(0), fw = R(0), fe = R(0);
    const R rc = f(QR, 0), ac = f(QA, 0), hc = f(QH, 0);
    const R uc = f(QU, 0), vc = f(QV, 0), wc = f(QW, 0);
    // on offset storage rho_a + rho_b = rho'_a + rho'_b + 2 rho0, and h is h'
    // (the h0 part of the energy flux is h0 times the mass flux, added below)
    const R r2 = R(2.0 * P.rho0);
#pragma unroll
    for (int l = S; l >= 1; l--) {
        const R c = R(FD<S>::a(l));
        const R rsp = Off<R>::on ? (rc + f(QR, l)) + r2 : rc + f(QR, l);
        const R rsm = Off<R>::on ? (f(QR, -l) + rc) + r2 : f(QR, -l) + rc;
        const R qp = rsp * (ac + f(QA, l));
        const R qm = rsm * (f(QA, -l) + ac);
        fr += c * (qp - qm);
        fu += c * (qp * (uc + f(QU, l)) - qm * (f(QU, -l) + uc));
        fv += c * (qp * (vc + f(QV, l)) - qm * (f(QV, -l) + vc));
        fw += c * (qp * (wc + f(QW, l)) - qm * (f(QW, -l) + wc));
        fe += c * (qp * (hc + f(QH, l)) - qm * (f(QH, -l) + hc));
    }
    const R sq = R(-0.5) * xs, sc = R(-0.25) * xs;
    out[0] = sq * fr;
    out[1] = sc * fu;
    out[2] = sc * fv;
    out[3] = sc * fw;
    out[4] = Off<R>::on ? sc * fe + R(P.h0) * out[0] : sc * fe;
}

// Inverse spacing of direction DIR at a point (the x metric in FP64 on
// offset storage).
template <typename R, int DIR, bool NONUNI>
__device__ __forceinline__ R dir_scale(const Params &P, double xpi) {
    const R d = R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    if (DIR == 0 && NONUNI) return Off<R>::on ? R(P.d_dx * xpi) : d * R(xpi);
    return d;
}

// The first derivative of one staged quantity along DIR at viscous order, as
// vel_stencils forms it.
template <typename R, int SV, int DIR, bool NONUNI, class Acc>
__device__ __forceinline__ R d1_of(const Params &P, Acc f, int q, double xpi) {
    R a1 = R(0);
#pragma unroll
    for (int k = SV; k >= 1; k--) a1 += R(FD<SV>::a(k)) * (f(q, k) - f(q, -k));
    if (Off<R>::on && DIR == 0 && NONUNI) return R((double)a1 * (P.d_dx * xpi));
    a1 = a1 * R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    if (DIR == 0 && NONUNI) a1 = a1 * R(xpi);
    return a1;
}

// PQ takes the pressure from the staged quantity QP (p itself) instead of
// forming rho R T at each point.
template <typename R, int S, int SV, int DIR, bool NONUNI, bool UNITV, bool PQ = false, class Acc>
__device__ __forceinline__ DirA<R> dir_terms_a(const Params &P, Acc f, const double *vs, double xpi) {
    const R d = R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    const R dd = R(DIR == 0 ? P.d_d2x : (DIR == 1 ? P.d_d2y : P.d_d2z));
    const R uc = f(QU, 0), vc = f(QV, 0), wc = f(QW, 0);
    const R muc = f(QM, 0), dilc = f(QD, 0);
    const R lamc = muc * R(P.invPrEc);
    const R Rg = R(P.Rgas);
    // viscous-order first derivatives of mu, T, dil; second derivative of T
    // (for mu = T/Re the mu derivative is the T derivative over Re)
    R d1m = R(0), d1t = R(0), d1d = R(0), d2t = R(0);
#pragma unroll
    for (int k = SV; k >= 1; k--) {
        const R a = R(FD<SV>::a(k));
        const R tp = f(QT, k), tm = f(QT, -k);
        if (!UNITV) d1m += a * (f(QM, k) - f(QM, -k));
        d1t += a * (tp - tm);
        d1d += a * (f(QD, k) - f(QD, -k));
        d2t += R(FD<SV>::b(k)) * (tp + tm);
    }
    R xs = d;
    if (DIR == 0 && NONUNI) xs = Off<R>::on ? R(P.d_dx * xpi) : d * R(xpi);
    d1t = d1t * xs;
    d1m = UNITV ? d1t * R(P.invRe) : d1m * xs;
    d1d = d1d * xs;
    if (DIR == 0 && NONUNI) {
        if constexpr (Off<R>::on) {
            double b = 0.0;
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) b += vs[it] * (double)f(QT, it - SV);
            d2t = R(b);
        } else {
            d2t = R(0);
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) d2t += R(vs[it]) * f(QT, it - SV);
        }
    } else {
        d2t = (d2t + R(FD<SV>::b(0)) * f(QT, 0)) * dd;
    }
    // advective-order pressure gradient, p = rho R T
    R d1p = R(0);
    if constexpr (PQ) {
#pragma unroll
        for (int k = S; k >= 1; k--) d1p += R(FD<S>::a(k)) * (f(QP, k) - f(QP, -k));
    } else if constexpr (Off<R>::on) {
        // p' / R = rho' T + rho0 T' on offset storage
        const R T0 = R(P.T0), r0 = R(P.rho0);
#pragma unroll
        for (int k = S; k >= 1; k--) {
            const R pp = f(QR, k) * (T0 + f(QT, k)) + r0 * f(QT, k);
            const R pm = f(QR, -k) * (T0 + f(QT, -k)) + r0 * f(QT, -k);
            d1p += R(FD<S>::a(k)) * (pp - pm);
        }
    } else {
#pragma unroll
        for (int k = S; k >= 1; k--) d1p += R(FD<S>::a(k)) * (f(QR, k) * f(QT, k) - f(QR, -k) * f(QT, -k));
    }
    d1p = PQ ? d1p * xs : d1p * (Rg * xs);
    const R d1l = d1m * R(P.invPrEc);

    R fl[5];
    dir_flux<R, S, DIR>(P, f, xs, fl);
    const R cr = fl[0], cu = fl[1], cv = fl[2], cw = fl[3], ce = fl[4];
    return DirA<R>{uc, vc, wc, muc, dilc, lamc, d1m, d1t, d1d, d2t, d1p, d1l, cr, cu, cv, cw, ce};
}

// The stress part of one direction's RHS (velocity gradients g, viscous
// second derivatives d2v) added to the gradient-free part a; the five
// partial right-hand sides in the reference's term order.
template <typename R, int DIR>
__device__ __forceinline__ void dir_terms_b(const Params &P, const DirA<R> &a, const R g[9], const R d2v[3],
                                            R out[5]) {
    const R uc = a.uc, vc = a.vc, wc = a.wc, muc = a.muc, dilc = a.dilc, lamc = a.lamc;
    const R d1m = a.d1m, d1t = a.d1t, d1d = a.d1d, d2t = a.d2t, d1p = a.d1p, d1l = a.d1l;
    const R cr = a.cr, cu = a.cu, cv = a.cv, cw = a.cw, ce = a.ce;
    // gradient index map: 0 dudx 1 dvdx 2 dwdx 3 dudy 4 dvdy 5 dwdy 6 dudz 7 dvdz 8 dwdz
    R su, sv, sw, e;
    const R twothird = R(2. / 3.);
    if (DIR == 0) {
        su = (2 * g[0] - twothird * dilc);
        sv = (g[1] + g[3]);
        sw = (g[2] + g[6]);
        e = muc * (su * g[0] + sv * g[1] + sw * g[2]);
    } else if (DIR == 1) {
        su = (g[3] + g[1]);
        sv = (2 * g[4] - twothird * dilc);
        sw = (g[5] + g[7]);
        // src/cuda_rhs.cu:175 multiplies the zy stress by dv/dz; fix_y_work
        // selects the physically consistent dw/dy.
        e = muc * (su * g[3] + sv * g[4] + sw * (P.fix_y_work ? g[5] : g[7]));
    } else {
        su = (g[6] + g[2]);
        sv = (g[7] + g[5]);
        sw = (2 * g[8] - twothird * dilc);
        e = muc * (su * g[6] + sv * g[7] + sw * g[8]);
    }
    su *= d1m;
    sv *= d1m;
    sw *= d1m;
    su = su + d2v[0] * muc;
    sv = sv + d2v[1] * muc;
    sw = sw + d2v[2] * muc;
    e = e + uc * su + vc * sv + wc * sw;
    const R ac3 = muc * d1d * R(1.0 / 3.0);
    if (DIR == 0) {
        e = e + d2t * lamc;
        e = e + d1t * d1l;
        su = su + ac3 - d1p;
        e = e + ac3 * uc;
    } else if (DIR == 1) {
        sv = sv + ac3;
        e = e + ac3 * vc;
        sv = sv - d1p;
        e = e + d2t * lamc;
        e = e + d1t * d1l;
    } else {
        sw = sw + ac3;
        e = e + ac3 * wc;
        sw = sw - d1p;
        e = e + d2t * lamc;
        e = e + d1t * d1l;
    }
    out[0] = cr;
    out[1] = su + cu;
    out[2] = sv + cv;
    out[3] = sw + cw;
    out[4] = e + ce;
}

// Everything else of one direction's part of the RHS; returns the five
// partial right-hand sides in the reference's term order.
template <typename R, int S, int SV, int DIR, bool NONUNI, bool UNITV, bool PQ = false, class Acc>
__device__ __forceinline__ void dir_terms(const Params &P, Acc f, const double *vs, double xpi, const R g[9],
                                         const R d2v[3], R out[5]) {
    dir_terms_b<R, DIR>(P, dir_terms_a<R, S, SV, DIR, NONUNI, UNITV, PQ>(P, f, vs, xpi), g, d2v, out);
}

__device__ __forceinline__ void block_max_atomic(double v, unsigned long long *out, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; w++) m = fmax(m, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// Bits of the stage-update flags.
enum { F_Q = 1, F_RA = 2, F_RB = 4, F_ACC = 8, F_R = 16, F_SPONGE = 32, F_VISC = 64, F_COMP = 128, F_QLO = 256 };

// The role-split kernel's plane march is compiled once per common stage
// kind with its flags as constants, plus once with run-time flags for the
// rest.  Constant flags fold the update's branches away, so the update of
// one plane and the stencils of the next form one basic block the compiler
// can interleave (7.5% on the order-8 FP64 kernel, 1% on the order-6
// channel).  F_VISC stays a run-time bit (a select, not a branch) and is not
// part of the key.
template <int... FS>
struct FlagList {};
template <class F>
__device__ __forceinline__ void with_flags(int, F &f, FlagList<>) {
    f(std::integral_constant<int, -1>{});
}
template <class F, int F0, int... FR>
__device__ __forceinline__ void with_flags(int key, F &f, FlagList<F0, FR...>) {
    if (key == F0)
        f(std::integral_constant<int, F0>{});
    else
        with_flags(key, f, FlagList<FR...>{});
}
// Stage kinds per storage type: RK4 stages 1-3 and 4; low-storage RK3
// stages 1, 2 and 3 (Kutta's first two stages have the flags of the first
// two); with K3 the last Kutta stage.  FP32 storage adds the compensated
// sums (F_COMP) and the low part of the step result (F_QLO); a sponge
// (always on in the boundary layer) adds F_SPONGE.
template <typename R, bool K3, bool SPONGE>
struct CommonFlags {
    static constexpr int C = Off<R>::on ? F_COMP : 0, L = Off<R>::on ? F_QLO : 0, SP = SPONGE ? F_SPONGE : 0;
    using type = std::conditional_t<
        K3, FlagList<F_Q | F_RA | F_RB | C | L | SP>,
        FlagList<F_Q | F_ACC | C | SP, F_Q | C | L | SP, F_Q | F_R | C | L | SP, F_Q | F_RA | F_R | C | L | SP,
                 F_Q | F_RA | C | L | SP>>;
};

__device__ __forceinline__ long long uniform64(long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// Operands of one plane's stage update: s1 and a second operand, ra
// (low-storage RK3, Kutta) or s2 (RK4's accumulator), rb as a third in the
// last Kutta stage (K3), and on FP32 storage the low parts of the
// compensated sets.  An array a stage does not use aliases the stage input,
// so its loads are cache hits; all are loaded unconditionally.
template <typename R>
struct Ops {
    R e1[5], e2[5], e3[5], e1l[5], e2l[5];
};

template <typename R, bool K3, class Arr>
__device__ __forceinline__ void load_ops(Arr arr, int fl, long poff, long fs, unsigned coff, Ops<R> &o) {
    const R *p1 = arr(0, poff), *p2 = arr((K3 || (fl & F_RA)) ? 1 : 3, poff);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        o.e1[q] = ld_at(p1 + q * fs, coff);
        o.e2[q] = ld_at(p2 + q * fs, coff);
    }
    if constexpr (K3) {
        const R *p3 = arr(2, poff);
#pragma unroll
        for (int q = 0; q < 5; q++) o.e3[q] = ld_at(p3 + q * fs, coff);
    } else {
#pragma unroll
        for (int q = 0; q < 5; q++) o.e3[q] = R(0);
    }
    if constexpr (Off<R>::on) {
        const R *l1 = arr(7, poff), *l2 = arr(8, poff);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            o.e1l[q] = ld_at(l1 + q * fs, coff);
            o.e2l[q] = K3 ? R(0) : ld_at(l2 + q * fs, coff);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 5; q++) o.e1l[q] = o.e2l[q] = R(0);
    }
}

// The end of one point's stage (plane k, column i, byte offset coff in the
// plane): sponge (src/sponge.cu:31-41), the viscous CFL maximum from the
// point's mu, then q_out = s1 + dt (w_ra ra + w_rb rb + w_r rhs),
// acc_out = s2 + dt v_r rhs and r_out = rhs as the flags select.  With FP32
// storage the update is evaluated in FP64 on the compensated pair hi + lo and
// rounded back to hi and lo (F_COMP).
template <typename R, bool K3, class Arr>
__device__ __forceinline__ void stage_update(const Params &P, const StagePtrs<R> &A, Arr arr, int fl,
                                             const double *s_w, long poff, long fs, unsigned coff, int i, int k,
                                             double dtd, R rhs[5], const Ops<R> &o, double mu, double inv_dx2,
                                             double &vmax) {
    const R dt = R(dtd);
    if (fl & F_SPONGE) {
        const long i2 = i + (long)P.mx * k, rs = (long)P.mx * P.mz;
        const R *qc = A.qin + poff;
        if constexpr (Off<R>::on) {
            const double sg = P.spx[i] + P.spz[k];
            const double off[5] = {P.rho0, 0.0, 0.0, 0.0, P.E0};
#pragma unroll
            for (int q = 0; q < 5; q++)
                rhs[q] += R(sg * (P.ref[i2 + q * rs] - ((double)ld_at(qc + q * fs, coff) + off[q])));
        } else {
            const R sg = R(P.spx[i] + P.spz[k]);
#pragma unroll
            for (int q = 0; q < 5; q++) rhs[q] += sg * (R(P.ref[i2 + q * rs]) - ld_at(qc + q * fs, coff));
        }
    }
    {
        // a select rather than a branch, so the update stays one basic block
        const double v1 = mu * inv_dx2, v2 = mu * P.d_d2y, v3 = mu * P.d_d2z;
        const double vm = v1 >= (v2 >= v3 ? v2 : v3) ? v1 : (v2 >= v3 ? v2 : v3);
        vmax = ((fl & F_VISC) && !(vmax >= vm)) ? vm : vmax;
    }
    if constexpr (Off<R>::on) {
        // FP64 update of the compensated pair, rounded to hi + lo
        const bool comp = fl & F_COMP;
        if (fl & F_Q) {
            R *const qo = arr(4, poff);
            const bool wlo = fl & F_QLO;
            R *const qol = wlo ? arr(9, poff) : qo;
#pragma unroll
            for (int q = 0; q < 5; q++) {
                double inc = 0.0;
                if (fl & F_RA) inc += s_w[0] * (double)o.e2[q];
                if (K3 && (fl & F_RB)) inc += s_w[1] * (double)o.e3[q];
                inc += s_w[2] * (double)rhs[q];
                const double x = ((double)o.e1[q] + (comp ? (double)o.e1l[q] : 0.0)) + dtd * inc;
                const R hi = R(x);
                st_at(qo + q * fs, coff, hi);
                if (wlo) st_at(qol + q * fs, coff, R(x - (double)hi));
            }
        }
        if (fl & F_R) {
            R *const ro = arr(5, poff);
#pragma unroll
            for (int q = 0; q < 5; q++) st_at(ro + q * fs, coff, rhs[q]);
        }
        if (fl & F_ACC) {
            R *const ao = arr(6, poff);
            R *const aol = comp ? arr(10, poff) : ao;
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const double x =
                    ((double)o.e2[q] + (comp ? (double)o.e2l[q] : 0.0)) + dtd * (s_w[3] * (double)rhs[q]);
                const R hi = R(x);
                st_at(ao + q * fs, coff, hi);
                if (comp) st_at(aol + q * fs, coff, R(x - (double)hi));
            }
        }
    } else {
        if (fl & F_Q) {
            R *const qo = arr(4, poff);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                R inc = R(0);
                if (fl & F_RA) inc += R(s_w[0]) * o.e2[q];
                if (K3 && (fl & F_RB)) inc += R(s_w[1]) * o.e3[q];
                inc += R(s_w[2]) * rhs[q];
                st_at(qo + q * fs, coff, o.e1[q] + dt * inc);
            }
        }
        if (fl & F_R) {
            R *const ro = arr(5, poff);
#pragma unroll
            for (int q = 0; q < 5; q++) st_at(ro + q * fs, coff, rhs[q]);
        }
        if (fl & F_ACC) {
            R *const ao = arr(6, poff);
#pragma unroll
            for (int q = 0; q < 5; q++) st_at(ao + q * fs, coff, o.e2[q] + dt * (R(s_w[3]) * rhs[q]));
        }
    }
}

// Fused RHS + stage update.  A TX x TY workgroup owns a tile of x-y columns
// and marches along z over one chunk of planes:
//   - plane k of the tile plus an x/y halo of width S is staged as primitives
//     in LDS (double buffered, one barrier per plane) for the x and y
//     stencils; with TX = TY = 16 and S <= 4 every lane stages at most one
//     halo point, whose source offset and ghost rule are fixed per lane;
//   - each lane keeps a (2S+1)-plane register window of its own column for
//     the z stencils;
//   - global loads are issued one plane ahead (window feed, halo) or at the
//     top of the plane (stage-update operands), so their latency overlaps the
//     stencil arithmetic.
// With FP32 storage (R = float) the state is stored as offsets from the
// reference state and the update is evaluated in FP64 on the compensated
// pair hi + lo, rounded back to hi and lo (F_COMP).
template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool UNITV, bool K3>
__device__ __forceinline__ void rhs_tile(const Params &P, const StagePtrs<R> &A, const int bb) {
    constexpr int H = S, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    constexpr int NL = UNITV ? NQ - 1 : NQ;  // staged quantities (mu rebuilt from T when mu = T/Re)
    constexpr int NH = 2 * H * TY + 2 * H * TX;
    static_assert(NH <= NT, "one halo point per lane");
    constexpr unsigned ESZ = sizeof(R);
    __shared__ R sq[2][NL][LY * LX];
    __shared__ double red[NT / 64 + 1];
    // Update weights, flags and the update's arrays (as byte offsets from the
    // stage input) live in LDS and are read where they are used, so they do
    // not hold scalar registers across the z march.
    __shared__ double s_w[4];
    __shared__ long long s_off[11];
    __shared__ int s_flags;
    if (threadIdx.x == 0) {
        s_w[0] = P.w_ra;
        s_w[1] = P.w_rb;
        s_w[2] = P.w_r;
        s_w[3] = P.v_r;
        const char *q0 = reinterpret_cast<const char *>(A.qin);
        s_off[0] = reinterpret_cast<const char *>(A.s1) - q0;
        s_off[1] = reinterpret_cast<const char *>(A.ra) - q0;
        s_off[2] = reinterpret_cast<const char *>(A.rb) - q0;
        s_off[3] = reinterpret_cast<const char *>(A.s2) - q0;
        s_off[4] = reinterpret_cast<const char *>(A.qout) - q0;
        s_off[5] = reinterpret_cast<const char *>(A.rout) - q0;
        s_off[6] = reinterpret_cast<const char *>(A.accout) - q0;
        s_off[7] = reinterpret_cast<const char *>(A.s1lo) - q0;
        s_off[8] = reinterpret_cast<const char *>(A.s2lo) - q0;
        s_off[9] = reinterpret_cast<const char *>(A.qoutlo) - q0;
        s_off[10] = reinterpret_cast<const char *>(A.accoutlo) - q0;
        s_flags = (P.write_q ? F_Q : 0) | (P.use_ra ? F_RA : 0) | (P.use_rb ? F_RB : 0) |
                  (P.write_acc ? F_ACC : 0) | (P.write_r ? F_R : 0) | (P.sponge ? F_SPONGE : 0) |
                  (P.want_visc_max ? F_VISC : 0) | (Off<R>::on && P.comp ? F_COMP : 0) |
                  (Off<R>::on && P.comp && P.write_qlo ? F_QLO : 0);
    }
    // array of the update at byte offset o from the stage input, plane poff
    auto arr = [&](int which, long poff) -> R * {
        const long long o = uniform64(s_off[which]);
        return reinterpret_cast<R *>(const_cast<char *>(reinterpret_cast<const char *>(A.qin)) + o) + poff;
    };

    const int tile_x = bb % P.ntx;
    const int rest = bb / P.ntx;
    const int tile_y = rest % P.nty;
    const int chunk = rest / P.nty;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    const int z0 = chunk * P.zc;
    const int z1 = min(z0 + P.zc, P.mz);
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int i = x0 + tx, j = y0 + ty;
    const bool act = (i < P.mx) && (j < P.my);
    const int cc = (ty + H) * LX + (tx + H);
    const long mxy = (long)P.mx * P.my;
    const long fs = A.fs;

    // own column (lanes past the domain edge stage the ghost of their column)
    int cside;
    const unsigned coff = src_offset<WALLX>(P, i, j, ESZ, cside);
    // this lane's halo point: x-halo columns first, then y-halo rows
    const bool hact = tid < NH;
    int hlx, hly;
    if (tid < 2 * H * TY) {
        hly = H + tid / (2 * H);
        const int c2 = tid % (2 * H);
        hlx = c2 < H ? c2 : TX + c2;
    } else {
        const int s2 = hact ? tid - 2 * H * TY : 0;
        hlx = H + s2 % TX;
        const int rr = s2 / TX;
        hly = rr < H ? rr : TY + rr;
    }
    int hside;
    const unsigned hoff = src_offset<WALLX>(P, x0 + hlx - H, y0 + hly - H, ESZ, hside);
    const int hslot = hly * LX + hlx;

    double xpi = 1.0, inv_dx2 = 1.0;
    double vs[2 * SV + 1];
#pragma unroll
    for (int it = 0; it < 2 * SV + 1; it++) vs[it] = 0.0;
    if (act) {
        const double dxv = P.dxv[i];
        inv_dx2 = 1.0 / (dxv * dxv);
        if (NONUNI) {
            xpi = P.xp[i];
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) vs[it] = P.vsx[it * P.mx + i];
        }
    }

    R win[NL][W];
#pragma unroll
    for (int q = 0; q < NL; q++)
#pragma unroll
        for (int a = 0; a < W; a++) win[q][a] = R(0);
    auto push = [&](const R *t, int a) {
#pragma unroll
        for (int q = 0; q < NL; q++) win[q][a] = t[q];
    };
    const R invRe = R(P.invRe);
    const R T0 = Off<R>::on ? R(P.T0) : R(0);  // T = T0 + staged T on offset storage
    // Every lane loads unconditionally (lanes past the edge keep the window of
    // their ghost column, lanes without a halo point load their mirror of
    // one, prefetches past the chunk wrap): with no loads under branches the
    // compiler can count outstanding loads exactly and wait for just the
    // ones it needs, so the prefetches stay in flight.
#pragma unroll
    for (int a = 0; a < W - 1; a++) {
        Raw<R> x;
        load_raw(A, (long)wrap1(z0 - H + a, P.mz) * mxy, coff, x);
        R t[NQ];
        to_prim<R, WALLX, UNITV>(P, x, cside, t);
        push(t, a);
    }
    Raw<R> zr, hr;  // prefetched: window feed (plane k + H) and halo point (plane k)
    load_raw(A, (long)wrap1(z0 + H, P.mz) * mxy, coff, zr);
    load_raw(A, (long)z0 * mxy, hoff, hr);
    double vmax = 0.0;
    const double dtd = (P.write_q || P.write_acc) ? *P.dt : 0.0;
    const R dpdz = P.forcing ? R(*P.dpdz) : R(0);
    __syncthreads();  // update parameters in LDS
    const int fl = __builtin_amdgcn_readfirstlane(s_flags);

    // run-time flags (fl): compiled with constant flags, as the role-split
    // form is, this kernel spills and runs at half speed
    for (int k = z0; k < z1; ++k) {
        const int b = (k - z0) & 1;
        const long poff = (long)k * mxy;
        // Operands of this plane's update, issued before the window feed and
        // halo prefetches with FP32 storage at order 8, after the plane
        // barrier (shorter register lifetimes) otherwise, which measured
        // faster each way (FP32: 2% at order 8; 4% on the order-6 channel and
        // 3-6% on the 128^3 order-6 Taylor-Green case the other way).
        Ops<R> ops;
        auto load_update_operands = [&]() { load_ops<R, K3>(arr, fl, poff, fs, coff, ops); };
        constexpr bool early_ops = sizeof(R) == 4 && S == 4;
        if constexpr (early_ops) load_update_operands();
        __builtin_amdgcn_sched_barrier(0);
        {
            R t[NQ];
            to_prim<R, WALLX, UNITV>(P, zr, cside, t);
            push(t, W - 1);
#pragma unroll
            for (int q = 0; q < NL; q++) sq[b][q][cc] = win[q][H];
            load_raw(A, (long)wrap1(k + 1 + H, P.mz) * mxy, coff, zr);
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            R t[NQ];
            to_prim<R, WALLX, UNITV>(P, hr, hside, t);
            if (hact) {
#pragma unroll
                for (int q = 0; q < NL; q++) sq[b][q][hslot] = t[q];
            }
            load_raw(A, (long)wrap1(k + 1, P.mz) * mxy, hoff, hr);
        }
        __syncthreads();
        if constexpr (!early_ops) load_update_operands();

        if (act) {
            auto FX = [&](int q, int a) -> R {
                return (UNITV && q == QM) ? (Off<R>::on ? T0 + sq[b][QT][cc + a] : sq[b][QT][cc + a]) * invRe
                                          : sq[b][q][cc + a];
            };
            auto FY = [&](int q, int a) -> R {
                return (UNITV && q == QM)
                           ? (Off<R>::on ? T0 + sq[b][QT][cc + a * LX] : sq[b][QT][cc + a * LX]) * invRe
                           : sq[b][q][cc + a * LX];
            };
            auto FZ = [&](int q, int a) -> R {
                return (UNITV && q == QM) ? (Off<R>::on ? T0 + win[QT][H + a] : win[QT][H + a]) * invRe
                                          : win[q][H + a];
            };
            R d1x[3], d2x[3], d1y[3], d2y[3], d1z[3], d2z[3];
            vel_stencils<R, SV, 0, NONUNI>(P, FX, vs, xpi, d1x, d2x);
            vel_stencils<R, SV, 1, false>(P, FY, vs, xpi, d1y, d2y);
            vel_stencils<R, SV, 2, false>(P, FZ, vs, xpi, d1z, d2z);
            const R g[9] = {d1x[0], d1x[1], d1x[2], d1y[0], d1y[1], d1y[2], d1z[0], d1z[1], d1z[2]};
            // x part (=), y part (+=), z part (+=) with the forcing, as
            // src/cuda_rhs.cu:123-127,255-259,389-393 accumulate them
            R rhs[5], part[5];
            dir_terms<R, S, SV, 0, NONUNI, UNITV>(P, FX, vs, xpi, g, d2x, rhs);
            dir_terms<R, S, SV, 1, false, UNITV>(P, FY, vs, xpi, g, d2y, part);
#pragma unroll
            for (int q = 0; q < 5; q++) rhs[q] = rhs[q] + part[q];
            dir_terms<R, S, SV, 2, false, UNITV>(P, FZ, vs, xpi, g, d2z, part);
            const R wc = win[QW][H];
            rhs[0] = rhs[0] + part[0];
            rhs[1] = rhs[1] + part[1];
            rhs[2] = rhs[2] + part[2];
            rhs[3] = rhs[3] + (part[3] + dpdz);
            rhs[4] = rhs[4] + (part[4] + dpdz * wc);

            stage_update<R, K3>(P, A, arr, fl, s_w, poff, fs, coff, i, k, dtd, rhs, ops, (double)FZ(QM, 0), inv_dx2,
                                vmax);
        }
#pragma unroll
        for (int q = 0; q < NL; q++)
#pragma unroll
            for (int a = 0; a < W - 1; a++) win[q][a] = win[q][a + 1];
    }
    if (P.want_visc_max) block_max_atomic(vmax, P.visc_max_out, red);
}

// Role-split form of rhs_tile for 2 TX TY lanes, one barrier per plane.  The
// waves of the first half (plane role) evaluate the x and y parts from LDS
// and apply the stage update; the waves of the second half (column role)
// keep the z register window of the tile's columns, stage each plane (its
// interior and its x/y halo) and evaluate the z part.  Each role holds about
// half the registers of rhs_tile, so two waves share a SIMD and hide each
// other's latency.  Per plane k (b = parity of k):
//   column: push plane k+S into the window, stage plane k's interior and
//           one halo point per lane, du/dz and dv/dz of plane k -> LDS;
//   -- barrier --
//   plane:  update of plane k-1 (its z part came through LDS), x part and
//           y part of plane k;
//   column: z part of plane k -> LDS, window shift.
// The plane role is the critical path: the column role takes all the work
// that crosses no role boundary (the staging; measured: the plane role
// waited at the barrier 3% of its time and the column role 52% with the
// halo in the plane role, and moving it gained 4% at order 8 and 12% on
// the order-6 channel; moving y stresses or the update instead, which
// then cross through LDS, measured slower).  Each role takes the velocity
// derivatives it needs from LDS itself.  Exchange buffers and the staged
// planes are double buffered by plane parity.  Both roles run the same
// barriers; the role is uniform within a wave.  The sums follow rhs_tile:
// (x part + y part) + z part, as src/cuda_rhs.cu accumulates them.
// BL: boundary-layer ghosts (src/boundary_condition_x.h, _z.h with
// boundaryLayer).  The bottom x ghosts are staged by their halo lanes
// (adiabatic reflection, blowing); the top x ghosts extrapolate the tile's
// last nodes, so their lanes form them from LDS after the plane is staged,
// behind a second barrier (the host requires mx to be a multiple of TX, so
// the last tile holds those nodes).  The z window extrapolates at both
// ends and carries p as a quantity of its own (QP).  The host also keeps
// every z chunk at least S planes long, so no window start reaches past an
// end.
template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool UNITV, bool K3, bool BL = false>
__device__ __forceinline__ void rhs_tile_split(const Params &P, const StagePtrs<R> &A, const int bb) {
    constexpr int H = S, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    constexpr int NL = UNITV ? NQ - 1 : NQ;  // staged quantities (mu rebuilt from T when mu = T/Re)
    constexpr int NH = 2 * H * TY + 2 * H * TX;
    static_assert(NH <= NT, "one halo point per column-role lane");
    static_assert(NT % 64 == 0, "each role is made of whole waves");
    static_assert(!BL || (WALLX && sizeof(R) == 8), "boundary layer: FP64, non-periodic x");
    constexpr unsigned ESZ = sizeof(R);
    __shared__ R sq[2][NL][LY * LX];
    // column -> plane role: du/dz, dv/dz of plane k (used after the barrier
    // of plane k) and the z part of plane k (used after the barrier of
    // plane k+1)
    __shared__ R xgz[2][2][NT];
    __shared__ R xyz[2][5][NT];
    __shared__ double red[2 * NT / 64 + 1];
    __shared__ double s_w[4];
    __shared__ long long s_off[11];
    __shared__ int s_flags;
    if (threadIdx.x == 0) {
        s_w[0] = P.w_ra;
        s_w[1] = P.w_rb;
        s_w[2] = P.w_r;
        s_w[3] = P.v_r;
        const char *q0 = reinterpret_cast<const char *>(A.qin);
        s_off[0] = reinterpret_cast<const char *>(A.s1) - q0;
        s_off[1] = reinterpret_cast<const char *>(A.ra) - q0;
        s_off[2] = reinterpret_cast<const char *>(A.rb) - q0;
        s_off[3] = reinterpret_cast<const char *>(A.s2) - q0;
        s_off[4] = reinterpret_cast<const char *>(A.qout) - q0;
        s_off[5] = reinterpret_cast<const char *>(A.rout) - q0;
        s_off[6] = reinterpret_cast<const char *>(A.accout) - q0;
        s_off[7] = reinterpret_cast<const char *>(A.s1lo) - q0;
        s_off[8] = reinterpret_cast<const char *>(A.s2lo) - q0;
        s_off[9] = reinterpret_cast<const char *>(A.qoutlo) - q0;
        s_off[10] = reinterpret_cast<const char *>(A.accoutlo) - q0;
        s_flags = (P.write_q ? F_Q : 0) | (P.use_ra ? F_RA : 0) | (P.use_rb ? F_RB : 0) |
                  (P.write_acc ? F_ACC : 0) | (P.write_r ? F_R : 0) | (P.sponge ? F_SPONGE : 0) |
                  (P.want_visc_max ? F_VISC : 0) | (Off<R>::on && P.comp ? F_COMP : 0) |
                  (Off<R>::on && P.comp && P.write_qlo ? F_QLO : 0);
    }
    auto arr = [&](int which, long poff) -> R * {
        const long long o = uniform64(s_off[which]);
        return reinterpret_cast<R *>(const_cast<char *>(reinterpret_cast<const char *>(A.qin)) + o) + poff;
    };

    const int tile_x = bb % P.ntx;
    const int rest = bb / P.ntx;
    const int tile_y = rest % P.nty;
    const int chunk = rest / P.nty;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    const int z0 = chunk * P.zc;
    const int z1 = min(z0 + P.zc, P.mz);
    // the role is uniform per wave; reading it from the first lane lets the
    // compiler branch on a scalar
    const bool colrole = __builtin_amdgcn_readfirstlane((int)threadIdx.x) >= NT;
    const int tid = colrole ? (int)threadIdx.x - NT : (int)threadIdx.x;  // column of the tile
    const int tx = tid % TX, ty = tid / TX;
    const int i = x0 + tx, j = y0 + ty;
    const bool act = (i < P.mx) && (j < P.my);
    const int cc = (ty + H) * LX + (tx + H);
    const long mxy = (long)P.mx * P.my;
    const long fs = A.fs;
    int cside;
    const unsigned coff = src_offset<WALLX>(P, i, j, ESZ, cside);
    const R invRe = R(P.invRe);
    const R T0 = Off<R>::on ? R(P.T0) : R(0);  // T = T0 + staged T on offset storage
    double xpi = 1.0;
    if (NONUNI && act) xpi = P.xp[i];
    double vmax = 0.0;
    __syncthreads();  // update parameters in LDS
    const int fl = __builtin_amdgcn_readfirstlane(s_flags);
    // staged-plane accessors along x and y (plane parity bk)
    auto FXb = [&](int bk) {
        return [&, bk](int q, int a) -> R {
            return (UNITV && q == QM) ? (Off<R>::on ? T0 + sq[bk][QT][cc + a] : sq[bk][QT][cc + a]) * invRe
                                      : sq[bk][q][cc + a];
        };
    };
    auto FYb = [&](int bk) {
        return [&, bk](int q, int a) -> R {
            return (UNITV && q == QM) ? (Off<R>::on ? T0 + sq[bk][QT][cc + a * LX] : sq[bk][QT][cc + a * LX]) * invRe
                                      : sq[bk][q][cc + a * LX];
        };
    };

    // this lane's halo point: x-halo columns first, then y-halo rows (lanes
    // past the ring write an unread corner slot of the plane)
    const bool hact = tid < NH;
    int hlx, hly;
    if (tid < 2 * H * TY) {
        hly = H + tid / (2 * H);
        const int c2 = tid % (2 * H);
        hlx = c2 < H ? c2 : TX + c2;
    } else {
        const int s2 = hact ? tid - 2 * H * TY : 0;
        hlx = H + s2 % TX;
        const int rr = s2 / TX;
        hly = rr < H ? rr : TY + rr;
    }
    int hside;
    const unsigned hoff = src_offset<WALLX>(P, x0 + hlx - H, y0 + hly - H, ESZ, hside);
    const int hslot = hly * LX + hlx;
    Raw<R> hr;  // prefetched halo point (plane k), in the role that stages it
    // BL blowing: sin(omega1 t), sin(10 t) and cos(y/lambda) of this lane's
    // halo row (y = j / d_dy as src/perturbation.h:50 forms it)
    double bs1 = 0.0, bs2 = 0.0, bcos = 0.0;
    if constexpr (BL) {
        if (P.perturbed) {
            const double tn = *P.time;
            bs1 = sin(P.blow_om1 * tn);
            bs2 = sin(10.0 * tn);
            bcos = cos(((double)wrapi(y0 + hly - H, P.my) / P.d_dy) / P.blow_lam);
        }
    }
    // the halo point of plane k into LDS buffer b, then the prefetch of
    // plane k+1; BL: the source point's primitives, at the bottom the
    // adiabatic reflection with the blowing (the top is formed by top_ghost)
    auto stage_halo = [&](int k, int b) {
        R t[NQ];
        if constexpr (BL) {
            to_prim<R, false, UNITV>(P, hr, 0, t);
            if (hside < 0) {
                double ub = 0.0;
                const bool blow = P.perturbed && bl_blow(k, bs1, bs2, bcos, ub);
                bl_bottom_ghost<R>(P, blow, R(ub), t[QR], t[QU], t[QV], t[QW], t[QT], t[QH]);
            }
        } else {
            to_prim<R, WALLX, UNITV>(P, hr, hside, t);
        }
        const int ws = hact ? hslot : 0;
#pragma unroll
        for (int q = 0; q < NL; q++) sq[b][q][ws] = t[q];
        load_raw(A, (long)wrap1(k + 1, P.mz) * mxy, hoff, hr);
    };
    // BL: the top x ghost of the halo point from the staged plane (topBCxExt
    // on u, v, w, p, T, dil; rho, h, mu from the ghost p and T:
    // src/boundary_condition_x.h:60-74, src/boundary.h:121-148)
    auto top_ghost = [&](int b) {
        if (hact && hside > 0) {
            const int d = (x0 + hlx - H) - (P.mx - 1);
            const int ls = hly * LX + (P.mx - 1 - x0 + H), is = ls - d;
            auto ext = [&](int q) { return R(2) * sq[b][q][ls] - sq[b][q][is]; };
            const R Rg = R(P.Rgas);
            const R ug = ext(QU), vg = ext(QV), wg = ext(QW), tg = ext(QT), dg = ext(QD);
            const R pg = R(2) * (sq[b][QR][ls] * Rg * sq[b][QT][ls]) - sq[b][QR][is] * Rg * sq[b][QT][is];
            sq[b][QU][hslot] = ug;
            sq[b][QV][hslot] = vg;
            sq[b][QW][hslot] = wg;
            sq[b][QT][hslot] = tg;
            sq[b][QD][hslot] = dg;
            sq[b][QH][hslot] = tg * R(P.cpR) + R(0.5) * (ug * ug + vg * vg + wg * wg);
            sq[b][QR][hslot] = pg / (Rg * tg);
            if constexpr (!UNITV) sq[b][QM % NL][hslot] = R(visc_pow((double)tg, P.viscexp)) * R(P.invRe);
        }
    };

    if (colrole) {
        load_raw(A, (long)z0 * mxy, hoff, hr);
        constexpr int NW = BL ? NL + 1 : NL;  // window quantities (BL: p last)
        R win[NW][W];
#pragma unroll
        for (int q = 0; q < NW; q++)
#pragma unroll
            for (int a = 0; a < W; a++) win[q][a] = R(0);
        auto push = [&](const R *t, int a) {
#pragma unroll
            for (int q = 0; q < NL; q++) win[q][a] = t[q];
            if constexpr (BL) win[NL][a] = t[QR] * R(P.Rgas) * t[QT];
        };
        // unconditional loads (see rhs_tile)
#pragma unroll
        for (int a = 0; a < W - 1; a++) {
            Raw<R> x;
            load_raw(A, (long)wrap1(z0 - H + a, P.mz) * mxy, coff, x);
            R t[NQ];
            to_prim<R, WALLX, UNITV>(P, x, cside, t);
            push(t, a);
        }
        Raw<R> zr;  // prefetched window feed (plane k + H)
        load_raw(A, (long)wrap1(z0 + H, P.mz) * mxy, coff, zr);
        auto FZ = [&](int q, int a) -> R {
            if (BL && q == QP) return win[NW - 1][H + a];
            return (UNITV && q == QM) ? (Off<R>::on ? T0 + win[QT][H + a] : win[QT][H + a]) * invRe
                                      : win[q < NW ? q : 0][H + a];
        };
        for (int k = z0; k < z1; ++k) {
            const int b = (k - z0) & 1;
            R d1z[3], d2z[3];
            {
                if (BL && k + H >= P.mz) {
                    // top z ghost: f(mz-1+d) = 2 f(mz-1) - f(mz-1-d) (topBCzExt)
                    const int d = k + H - (P.mz - 1);
#pragma unroll
                    for (int dd = 1; dd <= H; dd++)
                        if (d == dd)
#pragma unroll
                            for (int q = 0; q < NW; q++) win[q][W - 1] = R(2) * win[q][W - 1 - dd] - win[q][W - 1 - 2 * dd];
                } else {
                    R t[NQ];
                    to_prim<R, WALLX, UNITV>(P, zr, cside, t);
                    push(t, W - 1);
                }
                if (BL && k == 0) {
                    // bottom z ghosts: f(-d) = 2 f(0) - f(d) (botBCzExt)
#pragma unroll
                    for (int q = 0; q < NW; q++)
#pragma unroll
                        for (int a = 0; a < H; a++) win[q][a] = R(2) * win[q][H] - win[q][2 * H - a];
                }
#pragma unroll
                for (int q = 0; q < NL; q++) sq[b][q][cc] = win[q][H];
                load_raw(A, (long)wrap1(k + 1 + H, P.mz) * mxy, coff, zr);
                stage_halo(k, b);
            }
            vel_stencils<R, SV, 2, false>(P, FZ, nullptr, 1.0, d1z, d2z);
            xgz[b][0][tid] = d1z[0];
            xgz[b][1][tid] = d1z[1];
            __syncthreads();  // plane k staged; du/dz, dv/dz of plane k
            if constexpr (BL) {
                top_ghost(b);
                __syncthreads();  // x-top ghosts of plane k
            }
            if (act) {
                const auto FX = FXb(b);
                const auto FY = FYb(b);
                R part[5];
                const R g[9] = {R(0), R(0), d1_of<R, SV, 0, NONUNI>(P, FX, QW, xpi), R(0), R(0),
                                d1_of<R, SV, 1, false>(P, FY, QW, xpi), d1z[0], d1z[1], d1z[2]};
                dir_terms<R, S, SV, 2, false, UNITV, BL>(P, FZ, nullptr, 1.0, g, d2z, part);
#pragma unroll
                for (int q = 0; q < 5; q++) xyz[b][q][tid] = part[q];
            }
#pragma unroll
            for (int q = 0; q < NW; q++)
#pragma unroll
                for (int a = 0; a < W - 1; a++) win[q][a] = win[q][a + 1];
        }
        __syncthreads();  // z part of the last plane
    } else {
        // the plane role carries the longer chain per plane (x and y parts,
        // the update), so its waves win issue ties against the column role's
        // (measured: 2.5-4% on the order-6 channel, neutral at order 8)
        __builtin_amdgcn_s_setprio(3);
        double inv_dx2 = 1.0;
        double vs[2 * SV + 1];
#pragma unroll
        for (int it = 0; it < 2 * SV + 1; it++) vs[it] = 0.0;
        if (act) {
            const double dxv = P.dxv[i];
            inv_dx2 = 1.0 / (dxv * dxv);
            if (NONUNI) {
#pragma unroll
                for (int it = 0; it < 2 * SV + 1; it++) vs[it] = P.vsx[it * P.mx + i];
            }
        }
        const double dtd = (P.write_q || P.write_acc) ? *P.dt : 0.0;
        const R dpdz = P.forcing ? R(*P.dpdz) : R(0);
        // the previous plane: x part, y part, update operands, centre w and mu
        R rx[5], fy[5], wc = R(0);
        double muc = 0.0;
        Ops<R> ops;
#pragma unroll
        for (int q = 0; q < 5; q++) rx[q] = fy[q] = R(0);
        // the update of plane kk (parity bk) from its x and y parts and its z
        // part in LDS
        auto march = [&](auto FLC) {
            constexpr int FC = decltype(FLC)::value;
            const int flx = FC >= 0 ? FC | (fl & F_VISC) : fl;  // constant but for F_VISC
            auto finish = [&](int kk, int bk) {
                R rhs[5];
#pragma unroll
                for (int q = 0; q < 5; q++) rhs[q] = rx[q] + fy[q];
                rhs[0] = rhs[0] + xyz[bk][0][tid];
                rhs[1] = rhs[1] + xyz[bk][1][tid];
                rhs[2] = rhs[2] + xyz[bk][2][tid];
                rhs[3] = rhs[3] + (xyz[bk][3][tid] + dpdz);
                rhs[4] = rhs[4] + (xyz[bk][4][tid] + dpdz * wc);
                stage_update<R, K3>(P, A, arr, flx, s_w, (long)kk * mxy, fs, coff, i, kk, dtd, rhs, ops, muc, inv_dx2,
                                    vmax);
            };
            // one plane; FIN: the update of the previous plane runs in it (all
            // but the first plane of the chunk, which is peeled so the loop body
            // has no branch between the update and the stencils)
            auto body = [&](int k, auto FIN) {
                constexpr bool fin = decltype(FIN)::value;
                const int b = (k - z0) & 1;
                __syncthreads();  // plane k staged; du/dz, dv/dz of plane k; plane k-1's z part
                if constexpr (BL) __syncthreads();  // x-top ghosts of plane k
                if (act) {
                    if constexpr (fin) finish(k - 1, b ^ 1);
                    const auto FX = FXb(b);
                    const auto FY = FYb(b);
                    R d1x[3], d2x[3];
                    vel_stencils<R, SV, 0, NONUNI>(P, FX, vs, xpi, d1x, d2x);
                    R d1y[3], d2y[3];
                    vel_stencils<R, SV, 1, false>(P, FY, vs, xpi, d1y, d2y);
                    const R g[9] = {d1x[0], d1x[1], d1x[2], d1y[0], d1y[1], d1y[2],
                                    xgz[b][0][tid], xgz[b][1][tid], R(0)};
                    dir_terms<R, S, SV, 0, NONUNI, UNITV>(P, FX, vs, xpi, g, d2x, rx);
                    dir_terms<R, S, SV, 1, false, UNITV>(P, FY, vs, xpi, g, d2y, fy);
                    wc = FX(QW, 0);
                    muc = (double)FX(QM, 0);
                    load_ops<R, K3>(arr, flx, (long)k * mxy, fs, coff, ops);
                }
            };
            if (z0 < z1) body(z0, std::false_type{});
            for (int k = z0 + 1; k < z1; ++k) body(k, std::true_type{});
            __syncthreads();  // z part of the last plane
            if (act && z1 > z0) finish(z1 - 1, (z1 - 1 - z0) & 1);
        };
        with_flags(fl & ~F_VISC, march, typename CommonFlags<R, K3, BL>::type{});
    }
    if (P.want_visc_max) block_max_atomic(vmax, P.visc_max_out, red);
}

// Kernel form per storage type: FP64 runs the role-split tile (two waves
// per SIMD, each role within 256 registers); FP32 runs rhs_tile, which fits
// 256 registers on its own (the role split measured slower there).
template <typename R, bool UNITV>
constexpr bool rhs_split() {
    return sizeof(R) == 8;
}
// Waves per SIMD the RHS kernel is compiled for: two in every form (at most
// 256 VGPRs; FP32 with mu = T/Re with a few scratch spills, which cost less
// than the latency the second wave hides).
template <typename R, bool UNITV>
constexpr int rhs_waves() {
    return 2;
}
// Lanes of one RHS workgroup for a TX x TY tile.
template <typename R, bool UNITV, int TX, int TY>
constexpr int rhs_lanes() {
    return (rhs_split<R, UNITV>() ? 2 : 1) * TX * TY;
}

template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool UNITV, bool K3, bool BL = false>
__global__ __launch_bounds__((rhs_lanes<R, UNITV, TX, TY>()), (rhs_waves<R, UNITV>())) void k_rhs(Params P,
                                                                                                   StagePtrs<R> A) {
    const int bb = remap_block(blockIdx.x, gridDim.x, P.xcd_remap);
    if constexpr (rhs_split<R, UNITV>()) {
        rhs_tile_split<R, S, SV, TX, TY, WALLX, NONUNI, UNITV, K3, BL>(P, A, bb);
    } else {
        static_assert(!BL, "the boundary layer runs the role-split (FP64) kernel");
        rhs_tile<R, S, SV, TX, TY, WALLX, NONUNI, UNITV, K3>(P, A, bb);
    }
}

// Dilatation of the stage input: du/dx + dv/dy + dw/dz at viscous order
// (src/calc_stress.cu:20-63, 87-96).  Same tiling and prefetch scheme as
// k_rhs with halo SV: u and v of plane k (tile + halo) in double-buffered
// LDS, w of the own column in a register window.
// BL: the boundary layer's velocity ghosts (BCxderVel, BCzderVel): u odd
// about the bottom face with the blowing inside its band, extrapolated about
// the last node at the top (formed from LDS behind a second barrier, as in
// k_rhs); w extrapolated about both end planes in the window.
template <typename R, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool BL = false>
__device__ __forceinline__ void dil_tile(const Params &P, const StagePtrs<R> &A, R *dil, const int bb) {
    constexpr int H = SV, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    // halo lanes: the x-halo points (which need u only) first, padded to
    // whole waves, then the y-halo points (v only), so that each wave loads
    // density and one momentum field, the same one for all its lanes
    constexpr int NX = 2 * H * TY, XPAD = (NX + 63) / 64 * 64, NY = 2 * H * TX;
    static_assert(XPAD + NY <= NT, "one halo point per lane");
    constexpr unsigned ESZ = sizeof(R);
    __shared__ R su[2][LY * LX], sv[2][LY * LX];
    const int tile_x = bb % P.ntx;
    const int rest = bb / P.ntx;
    const int tile_y = rest % P.nty;
    const int chunk = rest / P.nty;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    const int z0 = chunk * P.zc;
    const int z1 = min(z0 + P.zc, P.mz);
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int i = x0 + tx, j = y0 + ty;
    const bool act = (i < P.mx) && (j < P.my);
    const int cc = (ty + H) * LX + (tx + H);
    const long mxy = (long)P.mx * P.my;
    const long fs = A.fs;

    int cside;
    const unsigned coff = src_offset<WALLX>(P, i, j, ESZ, cside);
    const bool xh = __builtin_amdgcn_readfirstlane(tid) < XPAD;  // uniform per wave
    const bool hact = xh ? tid < NX : tid - XPAD < NY;
    int hlx, hly;
    if (xh) {
        const int t2 = hact ? tid : 0;
        hly = H + t2 / (2 * H);
        const int c2 = t2 % (2 * H);
        hlx = c2 < H ? c2 : TX + c2;
    } else {
        const int s2 = hact ? tid - XPAD : 0;
        hlx = H + s2 % TX;
        const int rr = s2 / TX;
        hly = rr < H ? rr : TY + rr;
    }
    const int hq = xh ? 1 : 2;  // momentum field of this wave's halo points
    int hside;
    const unsigned hoff = src_offset<WALLX>(P, x0 + hlx - H, y0 + hly - H, ESZ, hside);
    const int hslot = hly * LX + hlx;
    // velocities are antisymmetric across an isothermal no-slip wall
    const R csg = cside ? R(-1) : R(1), hsg = hside ? R(-1) : R(1);
    double bs1 = 0.0, bs2 = 0.0, bcos = 0.0;  // BL blowing (see rhs_tile_split)
    if constexpr (BL) {
        static_assert(WALLX && sizeof(R) == 8, "boundary layer: FP64, non-periodic x");
        if (P.perturbed) {
            const double tn = *P.time;
            bs1 = sin(P.blow_om1 * tn);
            bs2 = sin(10.0 * tn);
            bcos = cos(((double)wrapi(y0 + hly - H, P.my) / P.d_dy) / P.blow_lam);
        }
    }
    double xpi = 1.0;
    if (NONUNI && act) xpi = P.xp[i];
    const R r0 = Off<R>::on ? R(P.rho0) : R(0);  // rho = rho0 + stored value on offset storage
    auto rho = [&](R stored) -> R { return Off<R>::on ? r0 + stored : stored; };

    // the own column: rho and rho w of the window plane, rho u and rho v of
    // the next staged plane (its density is the window's); a halo point:
    // rho and this wave's momentum field
    auto ldw = [&](long poff, R v2[2]) {
        const R *b = A.qin + poff;
        v2[0] = ld_at(b, coff);
        v2[1] = ld_at(b + 3 * fs, coff);
    };
    auto ldc = [&](long poff, R v2[2]) {
        const R *b = A.qin + poff;
        v2[0] = ld_at(b + fs, coff);
        v2[1] = ld_at(b + 2 * fs, coff);
    };
    auto ldh = [&](long poff, R v2[2]) {
        const R *b = A.qin + poff;
        v2[0] = ld_at(b, hoff);
        v2[1] = ld_at(b + hq * fs, hoff);
    };
    // unconditional loads (see rhs_tile): lanes past the edge and lanes
    // without a halo point load valid points they do not use.  Window: w and
    // 1/rho of the own column for planes k-H..k+H.
    R ww[W], ri[W];
#pragma unroll
    for (int a = 0; a < W - 1; a++) {
        R t[2];
        ldw((long)wrap1(z0 - H + a, P.mz) * mxy, t);
        ri[a] = recip(rho(t[0]));
        ww[a] = t[1] * ri[a];
    }
    ww[W - 1] = ri[W - 1] = R(0);
    R zr[2], cr[2], hr[2];
    ldw((long)wrap1(z0 + H, P.mz) * mxy, zr);
    ldc((long)z0 * mxy, cr);
    ldh((long)z0 * mxy, hr);

    for (int k = z0; k < z1; ++k) {
        const int b = (k - z0) & 1;
        const long poff = (long)k * mxy;
        const long pnext = (long)wrap1(k + 1, P.mz) * mxy;
        if (BL && k + H >= P.mz) {
            const int d = k + H - (P.mz - 1);  // top z ghost (topBCzExt)
#pragma unroll
            for (int dd = 1; dd <= H; dd++)
                if (d == dd) ww[W - 1] = R(2) * ww[W - 1 - dd] - ww[W - 1 - 2 * dd];
        } else {
            ri[W - 1] = recip(rho(zr[0]));
            ww[W - 1] = zr[1] * ri[W - 1];
        }
        if (BL && k == 0) {
#pragma unroll
            for (int a = 0; a < H; a++) ww[a] = R(2) * ww[H] - ww[2 * H - a];  // bottom z ghosts (botBCzExt)
        }
        ldw((long)wrap1(k + 1 + H, P.mz) * mxy, zr);
        su[b][cc] = csg * (cr[0] * ri[H]);
        sv[b][cc] = csg * (cr[1] * ri[H]);
        ldc(pnext, cr);
        {
            const R vh = hsg * (hr[1] * recip(rho(hr[0])));
            if (hact) {
                if (xh) {
                    R uh = vh;
                    if constexpr (BL) {
                        double ub = 0.0;
                        if (hside < 0 && P.perturbed && bl_blow(k, bs1, bs2, bcos, ub)) uh = R(ub);
                    }
                    su[b][hslot] = uh;
                } else {
                    sv[b][hslot] = vh;
                }
            }
        }
        ldh(pnext, hr);
        __syncthreads();
        if constexpr (BL) {
            if (hact && xh && hside > 0) {
                const int d = (x0 + hlx - H) - (P.mx - 1);  // top x ghost of u (topBCxExt)
                const int ls = hly * LX + (P.mx - 1 - x0 + H);
                su[b][hslot] = R(2) * su[b][ls] - su[b][ls - d];
            }
            __syncthreads();
        }
        if (act) {
            R a1 = R(0), a2 = R(0), a3 = R(0);
#pragma unroll
            for (int kk = SV; kk >= 1; kk--) {
                const R c = R(FD<SV>::a(kk));
                a1 += c * (su[b][cc + kk] - su[b][cc - kk]);
                a2 += c * (sv[b][cc + kk * LX] - sv[b][cc - kk * LX]);
                a3 += c * (ww[H + kk] - ww[H - kk]);
            }
            if (Off<R>::on && NONUNI) {
                a1 = R((double)a1 * (P.d_dx * xpi));  // metric in FP64
            } else {
                a1 = a1 * R(P.d_dx);
                if (NONUNI) a1 = a1 * R(xpi);
            }
            a2 = a2 * R(P.d_dy);
            a3 = a3 * R(P.d_dz);
            st_at(dil + poff, coff, a1 + a2 + a3);
        }
#pragma unroll
        for (int a = 0; a < W - 1; a++) {
            ww[a] = ww[a + 1];
            ri[a] = ri[a + 1];
        }
    }
}

template <typename R, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool BL = false>
__global__ __launch_bounds__(TX *TY) void k_dil(Params P, StagePtrs<R> A, R *dil) {
    dil_tile<R, SV, TX, TY, WALLX, NONUNI, BL>(P, A, dil, remap_block(blockIdx.x, gridDim.x, P.xcd_remap));
}

}  // namespace cns
