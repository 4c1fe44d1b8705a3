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
// (src/cuda_derivs.h:192-224 for x, 248-268 for y, 292-312 for z).
template <typename R, int SV, int DIR, bool NONUNI, class Acc>
__device__ __forceinline__ void vel_stencils(const Params &P, Acc f, const double *vs, double xpi, R d1[3],
                                             R d2[3]) {
    const R d = R(DIR == 0 ? P.d_dx : (DIR == 1 ? P.d_dy : P.d_dz));
    const R dd = R(DIR == 0 ? P.d_d2x : (DIR == 1 ? P.d_d2y : P.d_d2z));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int q = QU + c;
        R a1 = R(0);
#pragma unroll
        for (int it = 0; it < SV; it++) {
            if (DIR == 0)
                a1 += R(P.cVF[it]) * (f(q, it - SV) - f(q, SV - it));
            else
                a1 += R(P.cVF[it]) * (f(q, it - SV) - f(q, SV - it)) * d;
        }
        if (DIR == 0) {
            a1 = a1 * d;
            if (NONUNI) a1 = a1 * R(xpi);
        }
        d1[c] = a1;
        R a2;
        if (DIR == 0 && NONUNI) {
            a2 = R(0);
#pragma unroll
            for (int it = 0; it < 2 * SV + 1; it++) a2 += R(vs[it]) * f(q, it - SV);
        } else {
            a2 = R(P.cVS[SV]) * f(q, 0) * dd;
#pragma unroll
            for (int it = 0; it < SV; it++) a2 += R(P.cVS[it]) * (f(q, it - SV) + f(q, SV - it)) * dd;
        }
        d2[c] = a2;
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
    R d1m = R(0), d1t = R(0), d1d = R(0), d2t;
#pragma unroll
    for (int it = 0; it < SV; it++) {
        const R c = R(P.cVF[it]);
        if (DIR == 0) {
            d1m += c * (f(QM, it - SV) - f(QM, SV - it));
            d1t += c * (f(QT, it - SV) - f(QT, SV - it));
            d1d += c * (f(QD, it - SV) - f(QD, SV - it));
        } else {
            d1m += c * (f(QM, it - SV) - f(QM, SV - it)) * d;
            d1t += c * (f(QT, it - SV) - f(QT, SV - it)) * d;
            d1d += c * (f(QD, it - SV) - f(QD, SV - it)) * d;
        }
    }
    if (DIR == 0) {
        d1m = d1m * d;
        d1t = d1t * d;
        d1d = d1d * d;
        if (NONUNI) {
            d1m = d1m * R(xpi);
            d1t = d1t * R(xpi);
            d1d = d1d * R(xpi);
        }
    }
    if (DIR == 0 && NONUNI) {
        d2t = R(0);
#pragma unroll
        for (int it = 0; it < 2 * SV + 1; it++) d2t += R(vs[it]) * f(QT, it - SV);
    } else {
        d2t = R(P.cVS[SV]) * f(QT, 0) * dd;
#pragma unroll
        for (int it = 0; it < SV; it++) d2t += R(P.cVS[it]) * (f(QT, it - SV) + f(QT, SV - it)) * dd;
    }
    // advective-order pressure gradient, p = rho R T
    R d1p = R(0);
#pragma unroll
    for (int it = 0; it < S; it++) {
        const R pm = f(QR, it - S) * Rg * f(QT, it - S);
        const R pp = f(QR, S - it) * Rg * f(QT, S - it);
        if (DIR == 0)
            d1p += R(P.cF[it]) * (pm - pp);
        else
            d1p += R(P.cF[it]) * (pm - pp) * d;
    }
    if (DIR == 0) {
        d1p = d1p * d;
        if (NONUNI) d1p = d1p * R(xpi);
    }
    const R d1l = d1m * R(P.invPrEc);

    // split-form convective fluxes, telescoped pair form
    const int QA = (DIR == 0) ? QU : (DIR == 1 ? QV : QW);
    R fr = R(0), fu = R(0), fv = R(0), fw = R(0), fe = R(0);
    const R rc = f(QR, 0), ac = f(QA, 0), hc = f(QH, 0);
#pragma unroll
    for (int l = 1; l <= S; l++) {
        const R c = R(P.cF[S - l]);
        // pair (i, i+l)
        R rs = rc + f(QR, l), as = ac + f(QA, l);
        R qp = c * rs * as;
        R pu = qp * (uc + f(QU, l)), pv = qp * (vc + f(QV, l)), pw = qp * (wc + f(QW, l));
        R pe = qp * (hc + f(QH, l));
        // pair (i-l, i)
        R rs2 = f(QR, -l) + rc, as2 = f(QA, -l) + ac;
        R qm = c * rs2 * as2;
        fr += qp - qm;
        fu += pu - qm * (f(QU, -l) + uc);
        fv += pv - qm * (f(QV, -l) + vc);
        fw += pw - qm * (f(QW, -l) + wc);
        fe += pe - qm * (f(QH, -l) + hc);
    }
    R sq = R(0.5) * d, sc = R(0.25) * d;
    R cr = sq * fr, cu = sc * fu, cv = sc * fv, cw = sc * fw, ce = sc * fe;
    if (DIR == 0 && NONUNI) {
        cr = cr * R(xpi);
        cu = cu * R(xpi);
        cv = cv * R(xpi);
        cw = cw * R(xpi);
        ce = ce * R(xpi);
    }

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
    const R ac3 = muc * d1d / R(3.0);
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

template <typename R, int S, int SV, int TX, int TY, bool WALLX, bool NONUNI>
__global__ __launch_bounds__(TX *TY) void k_rhs(Params P, StagePtrs<R> A) {
    constexpr int H = S, LX = TX + 2 * H, LY = TY + 2 * H, NT = TX * TY, W = 2 * H + 1;
    constexpr int NH = 2 * H * TY + 2 * H * TX;
    __shared__ R sq[NQ][LY * LX];
    __shared__ double red[NT / 64 + 1];

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

    R win[NQ][W];
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int a = 0; a < W; a++) win[q][a] = R(0);
    if (act) {
#pragma unroll
        for (int a = 0; a < W - 1; a++) {
            R t[NQ];
            load_prim<R, WALLX>(P, A, i, j, wrapi(z0 - H + a, P.mz), t);
#pragma unroll
            for (int q = 0; q < NQ; q++) win[q][a] = t[q];
        }
    }
    double vmax = 0.0;
    const R dt = P.write_q ? R(*P.dt) : R(0);
    const R dpdz = P.forcing ? R(*P.dpdz) : R(0);

    for (int k = z0; k < z1; ++k) {
        if (act) {
            R t[NQ];
            load_prim<R, WALLX>(P, A, i, j, wrapi(k + H, P.mz), t);
#pragma unroll
            for (int q = 0; q < NQ; q++) win[q][W - 1] = t[q];
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < NQ; q++) sq[q][cc] = win[q][H];
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
            auto FZ = [&](int q, int a) -> R { return win[q][H + a]; };
            R d1x[3], d2x[3], d1y[3], d2y[3], d1z[3], d2z[3];
            vel_stencils<R, SV, 0, NONUNI>(P, FX, vs, xpi, d1x, d2x);
            vel_stencils<R, SV, 1, false>(P, FY, vs, xpi, d1y, d2y);
            vel_stencils<R, SV, 2, false>(P, FZ, vs, xpi, d1z, d2z);
            const R g[9] = {d1x[0], d1x[1], d1x[2], d1y[0], d1y[1], d1y[2], d1z[0], d1z[1], d1z[2]};
            R rx[5], ry[5], rz[5];
            dir_terms<R, S, SV, 0, NONUNI>(P, FX, vs, xpi, g, d2x, rx);
            dir_terms<R, S, SV, 1, false>(P, FY, vs, xpi, g, d2y, ry);
            dir_terms<R, S, SV, 2, false>(P, FZ, vs, xpi, g, d2z, rz);
            const R wc = win[QW][H];
            R rhs[5];
            rhs[0] = (rx[0] + ry[0]) + rz[0];
            rhs[1] = (rx[1] + ry[1]) + rz[1];
            rhs[2] = (rx[2] + ry[2]) + rz[2];
            rhs[3] = (rx[3] + ry[3]) + (rz[3] + dpdz);
            rhs[4] = (rx[4] + ry[4]) + (rz[4] + dpdz * wc);

            const long gidx = i + (long)P.mx * (j + (long)P.my * k);
            const long fs = A.fs;
            R qc[5];
#pragma unroll
            for (int q = 0; q < 5; q++) qc[q] = A.qin[gidx + q * fs];
            if (P.sponge) {
                const R sg = R(P.spx[i] + P.spz[k]);
                const long i2 = i + (long)P.mx * k, rs = (long)P.mx * P.mz;
#pragma unroll
                for (int q = 0; q < 5; q++) rhs[q] += sg * (R(P.ref[i2 + q * rs]) - qc[q]);
            }
            if (P.want_visc_max) {
                const double mu = (double)win[QM][H];
                const double v1 = mu / (dxvi * dxvi), v2 = mu * P.d_d2y, v3 = mu * P.d_d2z;
                const double vm = v1 >= (v2 >= v3 ? v2 : v3) ? v1 : (v2 >= v3 ? v2 : v3);
                vmax = vmax >= vm ? vmax : vm;
            }
            // stage update: straight-line per field, uniform flags only
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const long gq = gidx + q * fs;
                R base = R(P.w_qc) * qc[q];
                if (P.use_q0) base += R(P.w_q0) * A.q0[gq];
                R acc_old = R(0);
                if (P.use_acc) {
                    acc_old = A.acc[gq];
                    base += R(P.w_acc) * acc_old;
                }
                R inc = R(0);
                if (P.use_ra) inc += R(P.w_ra) * A.ra[gq];
                if (P.use_rb) inc += R(P.w_rb) * A.rb[gq];
                inc += R(P.w_r) * rhs[q];
                if (P.write_q) A.qout[gq] = base + dt * inc;
                if (P.write_r) A.rout[gq] = rhs[q];
                if (P.write_acc) A.acc[gq] = R(P.v_qc) * qc[q] + R(P.v_acc) * acc_old + dt * (R(P.v_r) * rhs[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; q++)
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
            for (int it = 0; it < SV; it++) {
                a1 += R(P.cVF[it]) * (su[cc + it - SV] - su[cc + SV - it]);
                a2 += R(P.cVF[it]) * (sv[cc + (it - SV) * LX] - sv[cc + (SV - it) * LX]) * R(P.d_dy);
                a3 += R(P.cVF[it]) * (ww[H + it - SV] - ww[H + SV - it]) * R(P.d_dz);
            }
            a1 = a1 * R(P.d_dx);
            if (NONUNI) a1 = a1 * R(xpi);
            dil[i + P.mx * (long)j + mxy * k] = a1 + a2 + a3;
        }
#pragma unroll
        for (int a = 0; a < W - 1; a++) ww[a] = ww[a + 1];
    }
}

}  // namespace cns
