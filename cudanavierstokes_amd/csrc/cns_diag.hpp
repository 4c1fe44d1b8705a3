// cns_diag.hpp -- on-device FP64 diagnostics of a state: the wall-normal
// channel profiles of calcAvgChan (src/init.cpp:150-208), the friction
// Reynolds number of printRes (src/init.cpp:210-256), and volume statistics
// (kinetic energy, enstrophy, mean squared dilatation) that the reference
// does not compute.
//
// All sums are formed in a fixed order (per-thread partials, then a fixed
// tree or a fixed sequential pass), so repeated calls give identical bits.
#pragma once
#include "cns_aux.hpp"

namespace cns {

// conserved variables of point g in FP64: stored value (+ low part) + offset
template <typename R>
__device__ __forceinline__ void full_state(const R *q, const R *lo, long g, long fs, double rho0, double E0,
                                           double c[5]) {
#pragma unroll
    for (int k = 0; k < 5; k++) c[k] = (double)q[g + k * fs] + (lo ? (double)lo[g + k * fs] : 0.0);
    c[0] += rho0;
    c[4] += E0;
}

// Per-x sums over a subset of the y-z rows (row, row + nch, ...): pass 0 sums
// rho, rho u, rho v, rho w, rho E; pass 1 sums the squared deviations of rho,
// u, v, w, rho E from the pass-0 means (mean[q*mx + i], q = 0..4: rho, Favre
// u, v, w, rho E).  part[(chunk*5 + q)*mx + i].
template <typename R>
__global__ void k_xprof_sums(int mx, long rows, long fs, const R *q, const R *lo, double rho0, double E0, int pass,
                             const double *mean, double *part, int nch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mx) return;
    double s[5] = {0, 0, 0, 0, 0};
    double m[5] = {0, 0, 0, 0, 0};
    if (pass == 1)
        for (int k = 0; k < 5; k++) m[k] = mean[k * mx + i];
    for (long row = blockIdx.y; row < rows; row += nch) {
        double c[5];
        full_state(q, lo, row * mx + i, fs, rho0, E0, c);
        if (pass == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) s[k] += c[k];
        } else {
            const double r = c[0];
            const double d[5] = {r - m[0], c[1] / r - m[1], c[2] / r - m[2], c[3] / r - m[3], c[4] - m[4]};
#pragma unroll
            for (int k = 0; k < 5; k++) s[k] += d[k] * d[k];
        }
    }
    for (int k = 0; k < 5; k++) part[((long)blockIdx.y * 5 + k) * mx + i] = s[k];
}

// Chunk partials -> averages over the my*mz rows (divided by my, then by mz,
// as the reference divides every term).  pass 0 writes the means (Favre for
// the velocities) to out[q*mx + i], pass 1 the mean squared deviations.
static __global__ void k_xprof_final(int mx, int my, int mz, const double *part, int nch, int pass, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mx) return;
    double s[5];
    for (int k = 0; k < 5; k++) {
        double a = 0.0;
        for (int c = 0; c < nch; c++) a += part[((long)c * 5 + k) * mx + i];
        s[k] = a / my / mz;
    }
    if (pass == 0) {
        out[0 * mx + i] = s[0];
        out[1 * mx + i] = s[1] / s[0];
        out[2 * mx + i] = s[2] / s[0];
        out[3 * mx + i] = s[3] / s[0];
        out[4 * mx + i] = s[4];
    } else {
        for (int k = 0; k < 5; k++) out[(5 + k) * mx + i] = s[k];
    }
}

// Tree sum of one value per thread into part[blockIdx.x] (RED_THREADS threads).
__device__ __forceinline__ void block_sum_to(double v, double *part) {
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int size = RED_THREADS / 2; size > 0; size /= 2) {
        if ((int)threadIdx.x < size) sh[threadIdx.x] += sh[threadIdx.x + size];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// printRes (src/init.cpp:210-240): per y-z row, dw/dx at node 0 from the
// advective-order stencil on w mirrored evenly about the bottom face
// (ub[s..2s] = w[0..s], ub[s-1-m] = w[m]), times xp[0]; u_tau =
// sqrt(mu_w |dw/dx| / rho) with mu_w = 1/Re (wall temperature 1), and the
// row's u_tau rho / mu_w.  part[blockIdx.x]: block sums over rows.
template <typename R, int S>
__global__ void k_ret(int mx, long rows, long fs, const R *q, const R *lo, double rho0, double E0, double dx,
                      double xp0, double muw, double c0, double c1, double c2, double c3, double *part) {
    const double cF[4] = {c0, c1, c2, c3};
    double acc = 0.0;
    for (long row = blockIdx.x * (long)blockDim.x + threadIdx.x; row < rows; row += (long)gridDim.x * blockDim.x) {
        double w[S + 1], r0 = 0.0;
#pragma unroll
        for (int m = 0; m <= S; m++) {
            double c[5];
            full_state(q, lo, row * mx + m, fs, rho0, E0, c);
            w[m] = c[3] / c[0];
            if (m == 0) r0 = c[0];
        }
        double ub[2 * S + 1];
#pragma unroll
        for (int m = S; m < 2 * S + 1; m++) ub[m] = w[m - S];
#pragma unroll
        for (int m = 0; m < S; m++) ub[m] = w[S - m - 1];
        double dudx = 0.0;
#pragma unroll
        for (int m = 0; m < S; m++) dudx += cF[m] * (ub[m] - ub[2 * S - m]) / dx;
        dudx *= xp0;
        const double ut = sqrt(muw * fabs(dudx) / r0);
        acc += ut * r0 / muw;
    }
    block_sum_to(acc, part);
}

// Volume statistics (no reference counterpart), weights dxv[i] dy dz:
//   part[0*RB + b]  sum rho
//   part[1*RB + b]  sum rho |u|^2 / 2
//   part[2*RB + b]  sum rho |omega|^2 / 2, omega = curl u
//   part[3*RB + b]  sum dil^2, dil = div u
// Velocity derivatives are the viscous-order central differences of the RHS
// (weights cVF, x derivative times xp on a stretched grid); across an
// isothermal wall the velocity ghost is the negated mirror value, as in the
// RHS's ghost rule; y, z (and x when periodic) wrap.
template <typename R, int SV>
__global__ void k_flow_stats(Params P, long fs, const R *q, const R *lo, int wallx, int nonuni, double *part) {
    const long mxy = (long)P.mx * P.my, n = mxy * P.mz;
    double s[4] = {0, 0, 0, 0};
    auto vel = [&](int i, int j, int k, double u[3]) {
        double sg = 1.0;
        if (wallx) {
            if (i < 0) {
                i = -1 - i;
                sg = -1.0;
            } else if (i >= P.mx) {
                i = 2 * P.mx - 1 - i;
                sg = -1.0;
            }
        } else {
            i = wrapi(i, P.mx);
        }
        j = wrapi(j, P.my);
        k = wrapi(k, P.mz);
        double c[5];
        full_state(q, lo, i + (long)P.mx * j + mxy * k, fs, P.rho0, P.E0, c);
        for (int a = 0; a < 3; a++) u[a] = sg * c[1 + a] / c[0];
    };
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        const int i = (int)(g % P.mx);
        const int j = (int)((g / P.mx) % P.my);
        const int k = (int)(g / mxy);
        double c[5];
        full_state(q, lo, g, fs, P.rho0, P.E0, c);
        const double r = c[0], u0 = c[1] / r, v0 = c[2] / r, w0 = c[3] / r;
        // grad[a][b] = d u_a / d x_b
        double grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int m = 1; m <= SV; m++) {
            const double a = P.cVF[SV - m];  // multiplies f[i-m] - f[i+m]
            double up[3], um[3];
            vel(i + m, j, k, up);
            vel(i - m, j, k, um);
            for (int b = 0; b < 3; b++) grad[b][0] += a * (um[b] - up[b]);
            vel(i, j + m, k, up);
            vel(i, j - m, k, um);
            for (int b = 0; b < 3; b++) grad[b][1] += a * (um[b] - up[b]);
            vel(i, j, k + m, up);
            vel(i, j, k - m, um);
            for (int b = 0; b < 3; b++) grad[b][2] += a * (um[b] - up[b]);
        }
        const double dxi = P.d_dx * (nonuni ? P.xp[i] : 1.0);
        for (int b = 0; b < 3; b++) {
            grad[b][0] *= dxi;
            grad[b][1] *= P.d_dy;
            grad[b][2] *= P.d_dz;
        }
        const double ox = grad[2][1] - grad[1][2], oy = grad[0][2] - grad[2][0], oz = grad[1][0] - grad[0][1];
        const double dil = grad[0][0] + grad[1][1] + grad[2][2];
        const double wgt = P.dxv[i] / P.d_dy / P.d_dz;
        s[0] += r * wgt;
        s[1] += 0.5 * r * (u0 * u0 + v0 * v0 + w0 * w0) * wgt;
        s[2] += 0.5 * r * (ox * ox + oy * oy + oz * oz) * wgt;
        s[3] += dil * dil * wgt;
    }
    __shared__ double sh[4][RED_THREADS];
    for (int a = 0; a < 4; a++) sh[a][threadIdx.x] = s[a];
    __syncthreads();
    for (int size = RED_THREADS / 2; size > 0; size /= 2) {
        if ((int)threadIdx.x < size)
            for (int a = 0; a < 4; a++) sh[a][threadIdx.x] += sh[a][threadIdx.x + size];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 4; a++) part[a * RED_BLOCKS + blockIdx.x] = sh[a][0];
}

// The statistics from the k_flow_stats partials, summed in block order by
// one thread (the order of the host's sum in cns_flow_stats, so both give
// the same bits): out[0..3] = kinetic energy, enstrophy, mean (div u)^2,
// mean density.
static __global__ void k_stats_final(const double *part, double vol, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < RED_BLOCKS; b++) sum[a] += part[a * RED_BLOCKS + b];
    out[0] = sum[1] / sum[0];
    out[1] = sum[2] / sum[0];
    out[2] = sum[3] / vol;
    out[3] = sum[0] / vol;
}

}  // namespace cns
