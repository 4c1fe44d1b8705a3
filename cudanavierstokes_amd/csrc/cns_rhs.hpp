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

namespace cns {

// Load one point of the stage input, convert to primitives and apply the
// x ghost rule.  (gi, gj) may lie outside the domain: y wraps periodically;
// x wraps (periodic) or mirrors to a wall ghost (src/boundary_condition_x.h).
template <typename R, bool WALLX>
__device__ __forceinline__ void load_prim(const Params &P, const StagePtrs<R> &A, int gi, int gj, int k,
                                          R *q) {
    gj = wrapi(gj, P.my);
    int side = 0;
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
    const long g = gi + (long)P.mx * (gj + (long)P.my * k);
    const long fs = A.fs;
    R r = A.qin[g];
    R u, v, w, t, h, mu;
    eos<R>(P, r, A.qin[g + fs], A.qin[g + 2 * fs], A.qin[g + 3 * fs], A.qin[g + 4 * fs], u, v, w, t, h, mu);
    R d = A.dil[g];
    if (WALLX && side != 0) wall_ghost<R>(P, side < 0 ? R(P.twall_bot) : R(P.twall_top), r, u, v, w, t, h, mu);
    q[QR] = r;
    q[QU] = u;
    q[QV] = v;
    q[QW] = w;
    q[QH] = h;
    q[QT] = t;
    q[QM] = mu;
    q[QD] = d;
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
        a1 = a1 * d;
        if (DIR == 0 && NONUNI) a1 = a1 * R(xpi);
        d1[c] = a1;
        if (DIR == 0 && NONUNI) {
            R b = R(0);
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) b += R(vs[it]) * f(q, it - SV);
            d2[c] = b;
        } else {
            d2[c] = (a2 + R(FD<SV>::b(0)) * f(q, 0)) * dd;
        }
    }
}

// Everything else of one direction's part of the RHS; returns the five
// partial right-hand sides in the reference's term order.
template <typename R, int S, int SV, int DIR, bool NONUNI, class Acc>
__device__ __forceinline__ void dir_terms(const Params &P, Acc f, const double *vs, double xpi, const R g[9],
                                         const R d2v[3], R out[5]) {
    const R d = R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    const R dd = R(DIR == 0 ? P.d_d2x : (DIR == 1 ? P.d_d2y : P.d_d2z));
    const R uc = f(QU, 0), vc = f(QV, 0), wc = f(QW, 0);
    const R muc = f(QM, 0), dilc = f(QD, 0);
    const R lamc = muc * R(P.invPrEc);
    const R Rg = R(P.Rgas);
    // viscous-order first derivatives of mu, T, dil; second derivative of T
    R d1m = R(0), d1t = R(0), d1d = R(0), d2t = R(0);
#pragma unroll
    for (int k = SV; k >= 1; k--) {
        const R a = R(FD<SV>::a(k));
        const R tp = f(QT, k), tm = f(QT, -k);
        d1m += a * (f(QM, k) - f(QM, -k));
        d1t += a * (tp - tm);
        d1d += a * (f(QD, k) - f(QD, -k));
        d2t += R(FD<SV>::b(k)) * (tp + tm);
    }
    R xs = d;
    if (DIR == 0 && NONUNI) xs = d * R(xpi);
    d1m = d1m * xs;
    d1t = d1t * xs;
    d1d = d1d * xs;
    if (DIR == 0 && NONUNI) {
        d2t = R(0);
#pragma unroll
        for (int it = 0; it < 2 * SV + 1; it++) d2t += R(vs[it]) * f(QT, it - SV);
    } else {
        d2t = (d2t + R(FD<SV>::b(0)) * f(QT, 0)) * dd;
    }
    // advective-order pressure gradient, p = rho R T
    R d1p = R(0);
#pragma unroll
    for (int k = S; k >= 1; k--) d1p += R(FD<S>::a(k)) * (f(QR, k) * f(QT, k) - f(QR, -k) * f(QT, -k));
    d1p = d1p * (Rg * xs);
    const R d1l = d1m * R(P.invPrEc);

    // split-form convective fluxes, telescoped pair form:
    // -d(fgh)/dx ~ -(d/4) sum_l a_l [(f_i+f_{i+l})(g_i+g_{i+l})(h_i+h_{i+l}) - (f_{i-l}+f_i)(...)(...)]
    const int QA = (DIR == 0) ? QU : (DIR == 1 ? QV : QW);
    R fr = R(0), fu = R(0), fv = R(0), fw = R(0), fe = R(0);
    const R rc = f(QR, 0), ac = f(QA, 0), hc = f(QH, 0);
#pragma unroll
    for (int l = S; l >= 1; l--) {
        const R c = R(FD<S>::a(l));
        const R qp = c * ((rc + f(QR, l)) * (ac + f(QA, l)));
        const R qm = c * ((f(QR, -l) + rc) * (f(QA, -l) + ac));
        fr += qp - qm;
        fu += qp * (uc + f(QU, l)) - qm * (f(QU, -l) + uc);
        fv += qp * (vc + f(QV, l)) - qm * (f(QV, -l) + vc);
        fw += qp * (wc + f(QW, l)) - qm * (f(QW, -l) + wc);
        fe += qp * (hc + f(QH, l)) - qm * (f(QH, -l) + hc);
    }
    const R sq = R(-0.5) * xs, sc = R(-0.25) * xs;
    const R cr = sq * fr, cu = sc * fu, cv = sc * fv, cw = sc * fw, ce = sc * fe;

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

template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI, bool UNITV>
__global__ __launch_bounds__(TX *TY) void k_rhs(Params P, StagePtrs<R> A) {
    constexpr int H = S, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    constexpr int NZ = UNITV ? 6 : 7;
    constexpr int NH = 2 * H * TY + 2 * H * TX;
    __shared__ R sq[NQ][LY * LX];
    __shared__ double red[NT / 64 + 1];
    // Stage-update weights, flags and output pointers live in LDS and are read
    // only by the update, so they do not occupy scalar registers across the
    // z march (the barriers in the loop keep these loads in place).
    __shared__ double s_w[9];
    __shared__ unsigned long long s_ptr[6];
    __shared__ int s_flags;
    if (threadIdx.x == 0) {
        s_w[0] = P.w_qc;
        s_w[1] = P.w_q0;
        s_w[2] = P.w_acc;
        s_w[3] = P.w_ra;
        s_w[4] = P.w_rb;
        s_w[5] = P.w_r;
        s_w[6] = P.v_qc;
        s_w[7] = P.v_acc;
        s_w[8] = P.v_r;
        s_ptr[0] = (unsigned long long)A.qout;
        s_ptr[1] = (unsigned long long)A.rout;
        s_ptr[2] = (unsigned long long)A.ra;
        s_ptr[3] = (unsigned long long)A.rb;
        s_ptr[4] = (unsigned long long)A.q0;
        s_ptr[5] = (unsigned long long)A.acc;
        s_flags = (P.use_q0 ? 1 : 0) | (P.use_acc ? 2 : 0) | (P.use_ra ? 4 : 0) | (P.use_rb ? 8 : 0) |
                  (P.write_q ? 16 : 0) | (P.write_r ? 32 : 0) | (P.write_acc ? 64 : 0) | (P.sponge ? 128 : 0) |
                  (P.want_visc_max ? 256 : 0);
    }

    const int bb = remap_block(blockIdx.x, gridDim.x, P.xcd_remap);
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

    double xpi = 1.0, dxvi = 1.0;
    double vs[2 * SV + 1];
#pragma unroll
    for (int it = 0; it < 2 * SV + 1; it++) vs[it] = 0.0;
    if (act) {
        dxvi = P.dxv[i];
        if (NONUNI) {
            xpi = P.xp[i];
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) vs[it] = P.vsx[it * P.mx + i];
        }
    }

    R win[NZ][W];
#pragma unroll
    for (int q = 0; q < NZ; q++)
#pragma unroll
        for (int a = 0; a < W; a++) win[q][a] = R(0);
    auto push = [&](const R *t, int a) {
        win[ZR][a] = t[QR];
        win[ZU][a] = t[QU];
        win[ZV][a] = t[QV];
        win[ZW][a] = t[QW];
        win[ZT][a] = t[QT];
        win[ZD][a] = t[QD];
        if (!UNITV) win[NZ - 1][a] = t[QM];
    };
    const R cpR = R(P.cpR), invRe = R(P.invRe);
    if (act) {
#pragma unroll
        for (int a = 0; a < W - 1; a++) {
            R t[NQ];
            load_prim<R, WALLX>(P, A, i, j, wrapi(z0 - H + a, P.mz), t);
            push(t, a);
        }
    }
    double vmax = 0.0;
    const R dt = P.write_q ? R(*P.dt) : R(0);
    const R dpdz = P.forcing ? R(*P.dpdz) : R(0);

    for (int k = z0; k < z1; ++k) {
        if (act) {
            R t[NQ];
            load_prim<R, WALLX>(P, A, i, j, wrapi(k + H, P.mz), t);
            push(t, W - 1);
        }
        // z-window accessor; h and (for mu = T/Re) mu are rebuilt on the fly
        auto FZ = [&](int q, int a) -> R {
            switch (q) {
            case QR: return win[ZR][H + a];
            case QU: return win[ZU][H + a];
            case QV: return win[ZV][H + a];
            case QW: return win[ZW][H + a];
            case QT: return win[ZT][H + a];
            case QD: return win[ZD][H + a];
            case QM: return UNITV ? win[ZT][H + a] * invRe : win[NZ - 1][H + a];
            default: {
                const R u = win[ZU][H + a], v = win[ZV][H + a], w = win[ZW][H + a];
                return win[ZT][H + a] * cpR + R(0.5) * (u * u + v * v + w * w);
            }
            }
        };
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < NQ; q++) sq[q][cc] = FZ(q, 0);
        } else {
            R t[NQ];
            load_prim<R, WALLX>(P, A, i, j, k, t);
#pragma unroll
            for (int q = 0; q < NQ; q++) sq[q][cc] = t[q];
        }
        for (int s = tid; s < NH; s += NT) {
            int lx, ly;
            if (s < 2 * H * TY) {
                ly = H + s / (2 * H);
                const int c2 = s % (2 * H);
                lx = c2 < H ? c2 : TX + c2;
            } else {
                const int s2 = s - 2 * H * TY;
                lx = H + s2 % TX;
                const int rr = s2 / TX;
                ly = rr < H ? rr : TY + rr;
            }
            R t[NQ];
            load_prim<R, WALLX>(P, A, x0 + lx - H, y0 + ly - H, k, t);
#pragma unroll
            for (int q = 0; q < NQ; q++) sq[q][ly * LX + lx] = t[q];
        }
        __syncthreads();

        if (act) {
            auto FX = [&](int q, int a) -> R { return sq[q][cc + a]; };
            auto FY = [&](int q, int a) -> R { return sq[q][cc + a * LX]; };
            // scheduling barriers between the phases keep the compiler from
            // hoisting every LDS read of the point to the top (register pressure)
            R d1x[3], d2x[3], d1y[3], d2y[3], d1z[3], d2z[3];
            vel_stencils<R, SV, 0, NONUNI>(P, FX, vs, xpi, d1x, d2x);
            __builtin_amdgcn_sched_barrier(0);
            vel_stencils<R, SV, 1, false>(P, FY, vs, xpi, d1y, d2y);
            __builtin_amdgcn_sched_barrier(0);
            vel_stencils<R, SV, 2, false>(P, FZ, vs, xpi, d1z, d2z);
            __builtin_amdgcn_sched_barrier(0);
            const R g[9] = {d1x[0], d1x[1], d1x[2], d1y[0], d1y[1], d1y[2], d1z[0], d1z[1], d1z[2]};
            // x part (=), y part (+=), z part (+=) with the forcing, as
            // src/cuda_rhs.cu:123-127,255-259,389-393 accumulate them
            R rhs[5], part[5];
            dir_terms<R, S, SV, 0, NONUNI>(P, FX, vs, xpi, g, d2x, rhs);
            __builtin_amdgcn_sched_barrier(0);
            dir_terms<R, S, SV, 1, false>(P, FY, vs, xpi, g, d2y, part);
#pragma unroll
            for (int q = 0; q < 5; q++) rhs[q] = rhs[q] + part[q];
            __builtin_amdgcn_sched_barrier(0);
            dir_terms<R, S, SV, 2, false>(P, FZ, vs, xpi, g, d2z, part);
            __builtin_amdgcn_sched_barrier(0);
            const R wc = win[ZW][H];
            rhs[0] = rhs[0] + part[0];
            rhs[1] = rhs[1] + part[1];
            rhs[2] = rhs[2] + part[2];
            rhs[3] = rhs[3] + (part[3] + dpdz);
            rhs[4] = rhs[4] + (part[4] + dpdz * wc);

            const long gidx = i + (long)P.mx * (j + (long)P.my * k);
            const long fs = A.fs;
            const int fl = __builtin_amdgcn_readfirstlane(s_flags);
            R qc[5];
#pragma unroll
            for (int q = 0; q < 5; q++) qc[q] = A.qin[gidx + q * fs];
            if (fl & 128) {
                const R sg = R(P.spx[i] + P.spz[k]);
                const long i2 = i + (long)P.mx * k, rs = (long)P.mx * P.mz;
#pragma unroll
                for (int q = 0; q < 5; q++) rhs[q] += sg * (R(P.ref[i2 + q * rs]) - qc[q]);
            }
            if (fl & 256) {
                const double mu = (double)FZ(QM, 0);
                const double v1 = mu / (dxvi * dxvi), v2 = mu * P.d_d2y, v3 = mu * P.d_d2z;
                const double vm = v1 >= (v2 >= v3 ? v2 : v3) ? v1 : (v2 >= v3 ? v2 : v3);
                vmax = vmax >= vm ? vmax : vm;
            }
            // stage update, one linear combination per field:
            //   q_out = w_qc q_in + w_q0 q0 + w_acc acc + dt (w_ra ra + w_rb rb + w_r rhs)
            R *const qout = (R *)s_ptr[0];
            R *const rout = (R *)s_ptr[1];
            const R *const ra = (const R *)s_ptr[2];
            const R *const rb = (const R *)s_ptr[3];
            const R *const q0 = (const R *)s_ptr[4];
            R *const acc = (R *)s_ptr[5];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const long gq = gidx + q * fs;
                R base = R(s_w[0]) * qc[q];
                if (fl & 1) base += R(s_w[1]) * q0[gq];
                R acc_old = R(0);
                if (fl & 2) {
                    acc_old = acc[gq];
                    base += R(s_w[2]) * acc_old;
                }
                R inc = R(0);
                if (fl & 4) inc += R(s_w[3]) * ra[gq];
                if (fl & 8) inc += R(s_w[4]) * rb[gq];
                inc += R(s_w[5]) * rhs[q];
                if (fl & 16) qout[gq] = base + dt * inc;
                if (fl & 32) rout[gq] = rhs[q];
                if (fl & 64) acc[gq] = R(s_w[6]) * qc[q] + R(s_w[7]) * acc_old + dt * (R(s_w[8]) * rhs[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < NZ; q++)
#pragma unroll
            for (int a = 0; a < W - 1; a++) win[q][a] = win[q][a + 1];
    }
    if (P.want_visc_max) block_max_atomic(vmax, P.visc_max_out, red);
}

// Dilatation of the stage input: du/dx + dv/dy + dw/dz at viscous order
// (src/calc_stress.cu:20-63, 87-96), same tiling as k_rhs with halo SV.
template <typename R, int SV, int TX, int TY, bool WALLX, bool NONUNI>
__global__ __launch_bounds__(TX *TY) void k_dil(Params P, StagePtrs<R> A, R *dil) {
    constexpr int H = SV, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    constexpr int NH = 2 * H * TY + 2 * H * TX;
    __shared__ R su[LY * LX], sv[LY * LX];
    const int bb = remap_block(blockIdx.x, gridDim.x, P.xcd_remap);
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
    double xpi = 1.0;
    if (NONUNI && act) xpi = P.xp[i];

    auto uv_at = [&](int gi, int gj, int k, R &u, R &v) {
        gj = wrapi(gj, P.my);
        bool ghost = false;
        if (WALLX) {
            if (gi < 0) {
                ghost = true;
                gi = -1 - gi;
            } else if (gi >= P.mx) {
                ghost = true;
                gi = 2 * P.mx - 1 - gi;
            }
            gi = gi < 0 ? 0 : (gi >= P.mx ? P.mx - 1 : gi);
        } else {
            gi = wrapi(gi, P.mx);
        }
        const long g = gi + P.mx * (long)gj + mxy * k;
        const R inv = R(1) / A.qin[g];
        u = A.qin[g + A.fs] * inv;
        v = A.qin[g + 2 * A.fs] * inv;
        if (ghost) {
            u = -u;
            v = -v;
        }
    };
    auto w_at = [&](int k) -> R {
        const long g = i + P.mx * (long)j + mxy * k;
        return A.qin[g + 3 * A.fs] / A.qin[g];
    };

    R ww[W];
    if (act) {
#pragma unroll
        for (int a = 0; a < W - 1; a++) ww[a] = w_at(wrapi(z0 - H + a, P.mz));
    }
    for (int k = z0; k < z1; ++k) {
        if (act) ww[W - 1] = w_at(wrapi(k + H, P.mz));
        __syncthreads();
        {
            R u, v;
            uv_at(i, j, k, u, v);
            su[cc] = u;
            sv[cc] = v;
        }
        for (int s = tid; s < NH; s += NT) {
            int lx, ly;
            if (s < 2 * H * TY) {
                ly = H + s / (2 * H);
                const int c2 = s % (2 * H);
                lx = c2 < H ? c2 : TX + c2;
            } else {
                const int s2 = s - 2 * H * TY;
                lx = H + s2 % TX;
                const int rr = s2 / TX;
                ly = rr < H ? rr : TY + rr;
            }
            R u, v;
            uv_at(x0 + lx - H, y0 + ly - H, k, u, v);
            su[ly * LX + lx] = u;
            sv[ly * LX + lx] = v;
        }
        __syncthreads();
        if (act) {
            R a1 = R(0), a2 = R(0), a3 = R(0);
#pragma unroll
            for (int k = SV; k >= 1; k--) {
                const R c = R(FD<SV>::a(k));
                a1 += c * (su[cc + k] - su[cc - k]);
                a2 += c * (sv[cc + k * LX] - sv[cc - k * LX]);
                a3 += c * (ww[H + k] - ww[H - k]);
            }
            a1 = a1 * R(P.d_dx);
            if (NONUNI) a1 = a1 * R(xpi);
            a2 = a2 * R(P.d_dy);
            a3 = a3 * R(P.d_dz);
            dil[i + P.mx * (long)j + mxy * k] = a1 + a2 + a3;
        }
#pragma unroll
        for (int a = 0; a < W - 1; a++) ww[a] = ww[a + 1];
    }
}

}  // namespace cns
